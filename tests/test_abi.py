"""The C-ABI library (CPU-side checks, no GPU compute): it loads, exports every entry point that
include/mcint.h declares, and its host-only calls behave."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mcint.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for must in ["mi_ctx_create", "mi_integrate", "mi_sigma_hat", "mi_vegas_adapt", "mi_unweight_compact",
                 "mi_adjoint", "mi_flow_train", "mi_rng_philox_fill", "mi_build_events", "mi_write_lhe"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    from cupythia_amd import _abi
    lib = C.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(_abi.PROTOTYPES), "ctypes prototypes out of sync with the header"


def test_struct_layouts_match_header():
    """sizeof of the ctypes mirrors == sizeof of the C structs (compiled here with gcc)."""
    import subprocess
    import tempfile
    from cupythia_amd import _abi as A
    src = '#include <stdio.h>\n#include "mcint.h"\nint main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",' \
          'sizeof(mi_physics),sizeof(mi_run),sizeof(mi_sigma_est),sizeof(mi_vegas_info),sizeof(mi_events),' \
          'sizeof(mi_sigma_grad),sizeof(mi_flow_info),offsetof(mi_run,n_points),sizeof(mi_particles),' \
          'offsetof(mi_particles,max_part),sizeof(mi_solve_opts),sizeof(mi_solve_info),' \
          'offsetof(mi_solve_info,final_points));return 0;}'
    src = src.replace('#include <stdio.h>', '#include <stdio.h>\n#include <stddef.h>')
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "t.c"), "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(d, "t"), os.path.join(d, "t.c")],
                       check=True)
        out = subprocess.run([os.path.join(d, "t")], check=True, capture_output=True, text=True).stdout.split()
    got = [C.sizeof(A.MiPhysics), C.sizeof(A.MiRun), C.sizeof(A.MiSigmaEst), C.sizeof(A.MiVegasInfo),
           C.sizeof(A.MiEvents), C.sizeof(A.MiSigmaGrad), C.sizeof(A.MiFlowInfo), A.MiRun.n_points.offset,
           C.sizeof(A.MiParticles), A.MiParticles.max_part.offset, C.sizeof(A.MiSolveOpts),
           C.sizeof(A.MiSolveInfo), A.MiSolveInfo.final_points.offset]
    assert [int(x) for x in out] == got


def test_host_only_entry_points():
    import cupythia_amd as mi
    from cupythia_amd import _abi
    lib = _abi.load()
    assert lib.mi_version() == _abi.MI_ABI_VERSION == 4
    p = mi.physics_default(1)
    assert p.pdf_source == mi.MI_PDF_TOY and p.fact_fac == 1.0 and tuple(p.mass_Q) == (1.5, 4.8)
    assert mi.PROC_CODES == (111, 112, 113, 114, 115, 116, 121, 122, 123, 124) and mi.MI_MAX_PROC == 10
    assert [cls._code for cls in mi.PROCESSES] == list(mi.PROC_CODES)
    assert p.mode == mi.MI_MODE_HADRONIC_3D and p.proc_mask == 0x1F and p.eCM == 13600.0
    assert p.pTHatMin == 20.0 and p.alphaS_value == 0.118 and p.alphaS_order == 1
    assert tuple(p.pdf_g) == (-0.1, 5.0) and tuple(p.pdf_val) == (0.5, 3.0) and tuple(p.pdf_sea) == (-0.2, 7.0)
    p0 = mi.physics_default(0)
    assert p0.mode == mi.MI_MODE_PARTONIC_1D and p0.sHat == 1e4 and p0.cos_max == 0.9
    assert lib.mi_physics_default(7, C.byref(_abi.MiPhysics())) == _abi.MI_ERR_ARG
    assert lib.mi_ctx_destroy(None) == 0
    assert lib.mi_last_error(None) == b"null context"


def test_no_gpu_context_creation_fails_cleanly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from cupythia_amd import _abi
    lib = _abi.load()
    h = C.c_void_p()
    assert lib.mi_ctx_create(0, C.byref(h)) == _abi.MI_ERR_HIP
    assert not h.value


def test_product_never_imports_the_oracle():
    """The product (cupythia_amd/) never imports, links or includes the checker."""
    pat = re.compile(r"(from|import)\s+oracle|libmcoracle|mcint_oracle|oracle_[a-z0-9]+\.(c|h)|oracle/_build")
    pkg = os.path.join(ROOT, "cupythia_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".cpp")) or f == "Makefile":
                assert not pat.search(open(os.path.join(dp, f)).read()), f


def _bare_context():
    """A Context whose wrappers can be called without a device: validation runs before any call
    into the library, so a bad argument raises before anything is launched (the library handle
    is absent, so a call that got past the checks would fail with AttributeError instead)."""
    import cupythia_amd as mi
    c = object.__new__(mi.Context)
    c.device, c.lib, c.h = 0, None, None
    return c


@pytest.mark.parametrize("case", [
    "flow_forward_flat", "flow_forward_f64", "flow_inverse_cols", "flow_backward_g_len", "sigma_hat_f32",
    "sigma_hat_len", "sigma_hat_id_int64", "reweight_pid_int64", "reweight_w_f32", "reweight_scales_len",
    "reweight_events_index_int32", "reweight_events_w_len", "events_index_dtype", "cpu_tensor", "noncontiguous",
    "not_a_tensor", "too_many_variations"])
def test_wrappers_reject_bad_arrays_before_any_launch(case):
    import torch
    import cupythia_amd as mi
    c = _bare_context()
    f64 = dict(dtype=torch.float64)
    n = 8
    a = torch.zeros(n, **f64)
    p = mi.physics_default(1)
    run = mi.make_run(p, n)
    i32 = torch.zeros(n, dtype=torch.int32)
    calls = {
        "flow_forward_flat": lambda: c.flow_forward(torch.zeros(3 * n, dtype=torch.float32)),
        "flow_forward_f64": lambda: c.flow_forward(torch.zeros((n, 3), **f64)),
        "flow_inverse_cols": lambda: c.flow_inverse(torch.zeros((n, 4), dtype=torch.float32)),
        "flow_backward_g_len": lambda: c.flow_backward(torch.zeros((n, 3), dtype=torch.float32),
                                                       torch.zeros(n + 1, dtype=torch.float32)),
        "sigma_hat_f32": lambda: c.sigma_hat(111, a.float(), a, a),
        "sigma_hat_len": lambda: c.sigma_hat(111, a, a[:4], a),
        "sigma_hat_id_int64": lambda: c.sigma_hat(114, a, a, a, id1=torch.zeros(n, dtype=torch.int64)),
        "reweight_pid_int64": lambda: c.reweight(p, a, a, a, a, torch.zeros(n, dtype=torch.int64), a, [(1, 1, 0)]),
        "reweight_w_f32": lambda: c.reweight(p, a, a, a, a, i32, a.float(), [(1, 1, 0)]),
        "reweight_scales_len": lambda: c.reweight_scales(p, a, a[:3], a, [1.0]),
        "reweight_events_index_int32": lambda: c.reweight_events(run, i32, i32, [(1, 1, 0)]),
        "reweight_events_w_len": lambda: c.reweight_events(run, torch.zeros(n, dtype=torch.int64), i32,
                                                           [(1, 1, 0)], w=a[:5]),
        "events_index_dtype": lambda: c.build_events(p, mi.Events(a, a, a, a, a, i32, i32, n, 1.0), key=1),
        "cpu_tensor": lambda: c.sigma_hat(111, a, a, a),
        "noncontiguous": lambda: c.sigma_hat(111, torch.zeros(2 * n, **f64)[::2], a, a),
        "not_a_tensor": lambda: c.sigma_hat(111, [1.0] * n, a, a),
        "too_many_variations": lambda: c.reweight_events(run, torch.zeros(n, dtype=torch.int64), i32,
                                                         [(1, 1, 0)] * 65),
    }
    with pytest.raises((ValueError, TypeError)) as e:
        calls[case]()
    if case == "cpu_tensor":
        assert "device tensor" in str(e.value)
    if case == "not_a_tensor":
        assert e.type is TypeError
