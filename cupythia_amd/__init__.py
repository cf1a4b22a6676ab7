"""cupythia_amd -- MI355X (gfx950) Monte-Carlo matrix-element integration.

Python face of the HIP library (include/mcint.h).  Device memory and streams come from
PyTorch (torch tensors on ``cuda``); every computation runs in the HIP kernels of
``_lib/libmcint.so``.  There is no CPU fallback: without the library every entry point raises.

The names mirror the reference interfaces the path replaces:
  * process identity: Pythia's ``Sigma2gg2gg`` ... ``Sigma2qqbar2qqbarNew`` and the heavy-quark
    pairs ``Sigma2gg2QQbar`` / ``Sigma2qqbar2QQbar`` (c cbar, b bbar) with ``code()``,
    ``name()``, ``inFlux()`` (pythia8317/include/Pythia8/SigmaQCD.h:188-392) and the batched
    ``sigmaHatWrap`` (pythia8317/include/Pythia8/SigmaProcess.h:516-520);
  * the sigma estimate: ``sigmaGen(i)``, ``sigmaErr(i)``, ``nTried``, ``nSelected``,
    ``nAccepted`` (pythia8317/include/Pythia8/Info.h:316-330,
    pythia8317/src/ProcessContainer.cc:1416-1472);
  * the counter RNG and the fused integration of cuPythia (cuPythia/common/rng.cuh:47-57,
    cuPythia/03_fused_resident_me/fused_gg2gg.cu:40-51).
"""
from ._abi import (  # noqa: F401
    LIB_PATH, LibraryMissing, MI_FLOW_NPARAMS, MI_GG2GG, MI_GG2QQBAR, MI_MAX_PROC,
    MI_MODE_HADRONIC_3D, MI_MODE_PARTONIC_1D, MI_PREC_FP32, MI_PREC_FP64, MI_PREC_FP32_HW, MI_QG2QG, MI_QQ2QQ,
    MI_QQBAR2GG, MI_QQBAR2QQBARNEW, MI_RNG_PHILOX, MI_RNG_SPLITMIX_COMPAT, MI_SAMPLER_FLAT,
    MI_SAMPLER_FLOW, MI_SAMPLER_VEGAS, MI_PDF_TOY, MI_PDF_GRID, MI_ABI_VERSION, MI_GG2CCBAR,
    MI_QQBAR2CCBAR, MI_GG2BBBAR, MI_QQBAR2BBBAR, PROC_CODES,
)
from .core import (  # noqa: F401
    Context, EventRecord, Events, MIError, SigmaEstimate, make_run, physics_default,
)
from .processes import (  # noqa: F401
    PROCESSES, Sigma2gg2gg, Sigma2gg2qqbar, Sigma2qg2qg, Sigma2qq2qq, Sigma2qqbar2gg,
    Sigma2qqbar2qqbarNew, Sigma2gg2ccbar, Sigma2qqbar2ccbar, Sigma2gg2bbbar, Sigma2qqbar2bbbar,
    process_by_code,
)
