// mi_pdf.hip -- proton PDFs from LHAPDF6 "lhagrid1" data files (Pythia's LHAGrid1): the reader,
// the device-resident tables in FP64 and FP32, and a batched evaluation entry point.
//
// Reader: pythia8317/src/PartonDistributions.cc:478-598 (LHAGrid1::init).  Header lines up to a
// line containing "---"; then per Q subgrid an x line (the same x nodes for every subgrid), a
// Q line, a flavour-id line, nx * nq_sub value lines (x outer, Q inner) and a "---" line.  The
// Q nodes of consecutive subgrids are concatenated; at each shared border both copies are
// replaced by their mean for the bisection (the logarithms keep the values as read), and the
// last Q of each subgrid is kept as the subgrid divider.  Flavour ids map to slots 0 = g (21 or
// 0), 1..5 = d u s c b, 6..10 = dbar ubar sbar cbar bbar, 11 = photon (22); others are dropped.
// Interpolation: GridPdf / grid_qw / grid_xf in mi_device.h (LHAGrid1::xfxevolve, :648-745).
#include "mi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mi {

struct PdfState {
    int nx = 0, nq = 0, nqSub = 0;
    std::vector<double> x, lnx, qk, lnq, val;        // host copies (FP64), val nq x nx x 12
    int nqSum[kGridMaxSub] = {};
    double qDiv[kGridMaxSub] = {};
    // device
    double *d64 = nullptr;    // xk | lnx | qk | lnq | val
    float *d32 = nullptr;     // val | ixinv | iqinv | xk | lnx | qk | lnq (val and the reciprocal
                              // tables 16-byte aligned: 12 floats per node / stencil start)
    unsigned short *dlut = nullptr;   // FP32 node guesses: x table | q table
    GridPdf<double> g64;      // views into d64
    GridPdf<float> g32;       // views into d32
    GridPdf<double> host64;   // views into the host copies
};

namespace {

int pfail(mi_ctx *ctx, int code, const std::string &msg)
{
    ctx->err = msg;
    return code;
}

int slot_of(int id)
{
    if (id == 21 || id == 0) return 0;
    if (id > 0 && id < 6) return id;
    if (id < 0 && id > -6) return 5 - id;
    if (id == 22) return 11;
    return -1;
}

// Parse the file into st (host part).  Returns MI_OK or an error with ctx->err set.
int parse_lhagrid1(mi_ctx *ctx, const char *path, PdfState &st)
{
    std::ifstream is(path);
    if (!is.good()) return pfail(ctx, MI_ERR_ARG, std::string("cannot open PDF grid file ") + path);
    std::string line;
    do {
        if (!std::getline(is, line)) return pfail(ctx, MI_ERR_ARG, "no '---' line in the PDF grid file");
    } while (line.find("---") == std::string::npos);
    std::vector<double> xs, qs;
    std::vector<std::vector<double>> blocks;         // per subgrid: nx x nq_sub x 12
    std::vector<int> starts;
    while (std::getline(is, line)) {
        std::istringstream isx(line);
        std::vector<double> xrow;
        double v;
        while (isx >> v) xrow.push_back(v);
        if (xrow.empty()) break;
        if ((int)starts.size() >= kGridMaxSub) return pfail(ctx, MI_ERR_RANGE, "more than 8 Q subgrids");
        if (starts.empty()) {
            xs = xrow;
        } else {
            if (xrow.size() != xs.size()) return pfail(ctx, MI_ERR_ARG, "a Q subgrid lists a different number of x nodes than the first");
            for (size_t i = 0; i < xs.size(); ++i)
                if (std::fabs(std::log(xrow[i]) - std::log(xs[i])) > 1e-5)
                    return pfail(ctx, MI_ERR_ARG, "the x nodes of a Q subgrid differ from the first subgrid's (ln x by more than 1e-5)");
        }
        if (!std::getline(is, line)) return pfail(ctx, MI_ERR_ARG, "truncated PDF grid file (Q line)");
        std::istringstream isq(line);
        const int q0 = (int)qs.size();
        while (isq >> v) qs.push_back(v);
        const int nqNow = (int)qs.size() - q0;
        if (nqNow < 2) return pfail(ctx, MI_ERR_ARG, "a Q subgrid needs at least 2 nodes");
        if (q0 > 0 && std::fabs(qs[q0] / qs[q0 - 1] - 1.0) > 1e-5)
            return pfail(ctx, MI_ERR_ARG, "a Q subgrid does not start at the last Q node of the previous one (relative 1e-5)");
        starts.push_back(q0);
        if (!std::getline(is, line)) return pfail(ctx, MI_ERR_ARG, "truncated PDF grid file (flavour line)");
        std::istringstream isid(line);
        std::vector<int> slots;
        int id;
        while (isid >> id) slots.push_back(slot_of(id));
        if (slots.empty()) return pfail(ctx, MI_ERR_ARG, "empty flavour line");
        std::vector<double> blk((size_t)xs.size() * nqNow * kGridNF, 0.0);
        for (size_t ix = 0; ix < xs.size(); ++ix)
            for (int iq = 0; iq < nqNow; ++iq) {
                if (!std::getline(is, line)) return pfail(ctx, MI_ERR_ARG, "truncated PDF grid file (values)");
                std::istringstream isv(line);
                double *dst = blk.data() + ((size_t)ix * nqNow + iq) * kGridNF;
                for (size_t k = 0; k < slots.size(); ++k) {
                    if (!(isv >> v)) return pfail(ctx, MI_ERR_ARG, "short PDF value line");
                    // the FP32 kernels keep a rounded copy: a value beyond the FP32 range would
                    // turn into inf there (and into NaN in the interpolation)
                    if (!(std::fabs(v) <= 3.4028234663852886e38))
                        return pfail(ctx, MI_ERR_RANGE, "PDF grid value " + std::to_string(v) +
                                                            " is not finite in FP32");
                    if (slots[k] >= 0) dst[slots[k]] = v;
                }
            }
        blocks.push_back(std::move(blk));
        if (!std::getline(is, line)) break;          // separator
    }
    if (starts.empty() || xs.size() < 4) return pfail(ctx, MI_ERR_ARG, "PDF grid needs >= 4 x nodes and one subgrid");
    st.nx = (int)xs.size();
    st.nq = (int)qs.size();
    st.nqSub = (int)starts.size();
    st.x = xs;
    st.lnx.resize(st.nx);
    for (int i = 0; i < st.nx; ++i) st.lnx[i] = std::log(xs[i]);
    st.qk = qs;
    st.lnq.resize(st.nq);
    for (int i = 0; i < st.nq; ++i) st.lnq[i] = std::log(qs[i]);
    for (int s = 0; s < st.nqSub; ++s) {
        const int end = (s + 1 < st.nqSub) ? starts[s + 1] : st.nq;
        st.nqSum[s] = end;
        st.qDiv[s] = qs[end - 1];
        if (s > 0) {
            const int b = starts[s];
            const double avg = 0.5 * (qs[b - 1] + qs[b]);
            st.qk[b - 1] = avg;
            st.qk[b] = avg;
        }
    }
    st.val.assign((size_t)st.nq * st.nx * kGridNF, 0.0);
    for (int s = 0; s < st.nqSub; ++s) {
        const int nqs = st.nqSum[s] - starts[s];
        for (int ix = 0; ix < st.nx; ++ix)
            for (int iq = 0; iq < nqs; ++iq)
                std::memcpy(st.val.data() + ((size_t)(starts[s] + iq) * st.nx + ix) * kGridNF,
                            blocks[s].data() + ((size_t)ix * nqs + iq) * kGridNF, sizeof(double) * kGridNF);
    }
    return MI_OK;
}

// Reciprocal table of 4-point Lagrange denominators for stencils starting at 0 .. n - 4:
// inv[12 m + k] = 1 / (t[m+i] - t[m+j]) over the (i, j != i) pairs in order (FP64, then
// rounded to FP32).
void lagrange_inv(const std::vector<double> &t, std::vector<float> &inv)
{
    const int n = (int)t.size();
    inv.assign((size_t)kGridNF * n, 0.0f);
    for (int m = 0; m + 3 < n; ++m) {
        int k = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (j != i) inv[(size_t)kGridNF * m + k++] = (float)(1.0 / (t[m + i] - t[m + j]));
    }
}

// Node guesses for a uniform binning of [lnt[0], lnt[n-1]] into nb bins: lut[j] = the last node
// (at most n - 2) whose log is at or below bin j's left edge.
constexpr int kLutX = 4096, kLutQ = 1024;
void node_lut(const std::vector<double> &lnt, int nb, std::vector<unsigned short> &lut, float &l0, float &ih)
{
    const int n = (int)lnt.size();
    const double a = lnt.front(), h = (lnt.back() - a) / nb;
    l0 = (float)a;
    ih = (float)(1.0 / h);
    int i = 0;
    for (int j = 0; j < nb; ++j) {
        const double e = a + j * h;
        while (i + 1 <= n - 2 && lnt[i + 1] <= e) ++i;
        lut.push_back((unsigned short)i);
    }
}

template <typename T>
void fill_view(GridPdf<T> &g, const PdfState &st)
{
    g.nx = st.nx; g.nq = st.nq; g.nqSub = st.nqSub;
    for (int s = 0; s < kGridMaxSub; ++s) { g.nqSum[s] = st.nqSum[s]; g.qDiv[s] = (T)st.qDiv[s]; }
    g.xMin = (T)st.x.front(); g.xMax = (T)st.x.back();
    g.qMin = (T)st.qk.front(); g.qMax = (T)st.qk.back();
    g.active = 0;
    g.factFac2 = (T)1;
}

__global__ void k_pdf_xf(GridPdf<double> g, uint64_t n, const double *__restrict__ x, const double *__restrict__ Q2,
                         double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = sqrt(Q2[i]);
    int m3q, n3q;
    double wq[4], f[11];
    grid_qw<double>(g, q, log(q), m3q, n3q, wq);
    grid_xf<double>(g, x[i], log(x[i]), m3q, n3q, wq, f);
    for (int k = 0; k < 11; ++k) out[i * kGridNF + k] = f[k];
    out[i * kGridNF + 11] = 0.0;
}

}  // namespace

void release_state(PdfState *&st)
{
    if (!st) return;
    (void)hipFree(st->d64);
    (void)hipFree(st->d32);
    (void)hipFree(st->dlut);
    delete st;
    st = nullptr;
}

void pdf_release(mi_ctx *ctx)
{
    if (!ctx) return;
    release_state(ctx->pdf);
    for (int s = 1; s < kPdfSlots; ++s) release_state(ctx->pdf_var[s]);
}

int pdf_slot_view(mi_ctx *ctx, int slot, GridPdf<double> *g64)
{
    PdfState *st = slot == 0 ? ctx->pdf : (slot > 0 && slot < kPdfSlots ? ctx->pdf_var[slot] : nullptr);
    if (slot < 0 || slot >= kPdfSlots) return pfail(ctx, MI_ERR_ARG, "PDF slot out of range");
    if (!st) return pfail(ctx, MI_ERR_STATE, "no PDF grid loaded in slot " + std::to_string(slot));
    *g64 = st->g64;
    g64->active = 1;
    return MI_OK;
}

int pdf_views(mi_ctx *ctx, const mi_physics &p, GridPdf<float> *g32, GridPdf<double> *g64)
{
    if (g32) *g32 = GridPdf<float>{};
    if (g64) *g64 = GridPdf<double>{};
    if (p.pdf_source != MI_PDF_GRID) return MI_OK;
    if (!ctx->pdf) return pfail(ctx, MI_ERR_STATE, "pdf_source = MI_PDF_GRID but no grid loaded (mi_pdf_load)");
    const double ff2 = p.fact_fac * p.fact_fac;
    if (g32) { *g32 = ctx->pdf->g32; g32->active = 1; g32->factFac2 = (float)ff2; }
    if (g64) { *g64 = ctx->pdf->g64; g64->active = 1; g64->factFac2 = ff2; }
    return MI_OK;
}

bool pdf_window(const mi_ctx *ctx, const mi_physics &p, size_t off, size_t max_bytes, GridPdf<float> *g,
                size_t *bytes)
{
    if (p.pdf_source != MI_PDF_GRID || !ctx->pdf || !(p.eCM > 0.0) || !(p.pTHatMin > 0.0)) return false;
    const PdfState &st = *ctx->pdf;
    const int nx = st.nx, nq = st.nq;
    // last node (FP32, as the device compares) at or below v, -1 if none
    auto last_le = [](const std::vector<double> &t, float v) {
        int i = -1;
        while (i + 1 < (int)t.size() && (float)t[i + 1] <= v) ++i;
        return i;
    };
    const float xlo = (float)(0.999 * 4.0 * p.pTHatMin * p.pTHatMin / (p.eCM * p.eCM));
    int x0 = 0;
    if (xlo > (float)st.x.front()) x0 = std::max(0, std::min(last_le(st.x, xlo), nx - 2) - 1);
    const float qlo = (float)(0.999 * p.fact_fac * p.pTHatMin), qhi = (float)(1.001 * p.fact_fac * 0.5 * p.eCM);
    const int r0 = qlo <= (float)st.qk.front() ? 0 : std::max(0, last_le(st.qk, qlo) - 3);
    const int r1 = qhi >= (float)st.qk.back() ? nq - 1 : std::min(nq - 1, last_le(st.qk, qhi) + 4);
    const size_t b = (size_t)(r1 - r0 + 1) * (size_t)(nx - x0) * kGridNF * sizeof(float);
    if (r1 < r0 || off % 16 != 0 || off + b > max_bytes) return false;
    g->woff = (int)off;
    g->wq0 = r0;
    g->wq1 = r1;
    g->wx0 = x0;
    g->wstride = nx - x0;
    *bytes = b;
    return true;
}

bool pdf_host_view(const mi_ctx *ctx, const mi_physics &p, GridPdf<double> *g)
{
    if (p.pdf_source != MI_PDF_GRID || !ctx->pdf) return false;
    *g = ctx->pdf->host64;
    g->active = 1;
    g->factFac2 = p.fact_fac * p.fact_fac;
    return true;
}

}  // namespace mi

using namespace mi;

extern "C" {

int mi_pdf_load(mi_ctx *ctx, const char *path) { return mi_pdf_load_slot(ctx, 0, path); }

int mi_pdf_load_slot(mi_ctx *ctx, int slot, const char *path)
{
    if (!ctx) return MI_ERR_ARG;
    if (!path) return pfail(ctx, MI_ERR_ARG, "path is NULL");
    if (slot < 0 || slot >= kPdfSlots) return pfail(ctx, MI_ERR_ARG, "PDF slot out of range");
    PdfState *st = new PdfState();
    int rc = parse_lhagrid1(ctx, path, *st);
    if (rc) { delete st; return rc; }
    hipError_t e = hipSetDevice(ctx->device);
    const size_t nx = st->nx, nq = st->nq, nv = (size_t)nq * nx * kGridNF;
    std::vector<float> ixinv, iqinv;
    lagrange_inv(st->lnx, ixinv);
    lagrange_inv(st->lnq, iqinv);
    // 2-3 point subgrids interpolate linearly: iqinv[12 m] = 1 / (lnq[m+1] - lnq[m])
    for (int s = 0; s < st->nqSub; ++s) {
        const int minS = s == 0 ? 0 : st->nqSum[s - 1], maxS = st->nqSum[s] - 1;
        if (maxS - minS < 3)
            for (int m = minS; m < maxS; ++m) iqinv[(size_t)kGridNF * m] = (float)(1.0 / (st->lnq[m + 1] - st->lnq[m]));
    }
    std::vector<double> h64;
    h64.reserve(2 * nx + 2 * nq + nv);
    h64.insert(h64.end(), st->x.begin(), st->x.end());
    h64.insert(h64.end(), st->lnx.begin(), st->lnx.end());
    h64.insert(h64.end(), st->qk.begin(), st->qk.end());
    h64.insert(h64.end(), st->lnq.begin(), st->lnq.end());
    h64.insert(h64.end(), st->val.begin(), st->val.end());
    // FP32 copy: rounded from FP64 (the FP32 logs are the FP64 logs rounded, not logs of the rounded
    // nodes), values first so that every node's 12 values are three aligned 16-byte words
    std::vector<float> h32;
    h32.reserve(nv + kGridNF * (nx + nq) + 2 * nx + 2 * nq);
    for (size_t i = 0; i < nv; ++i) h32.push_back((float)st->val[i]);
    h32.insert(h32.end(), ixinv.begin(), ixinv.end());
    h32.insert(h32.end(), iqinv.begin(), iqinv.end());
    for (double v : st->x) h32.push_back((float)v);
    for (double v : st->lnx) h32.push_back((float)v);
    for (double v : st->qk) h32.push_back((float)v);
    for (double v : st->lnq) h32.push_back((float)v);
    if (nx > 65535 || nq > 65535) { delete st; return pfail(ctx, MI_ERR_RANGE, "PDF grid above 65535 nodes"); }
    std::vector<unsigned short> lut;
    float xl0, xih, ql0, qih;
    node_lut(st->lnx, kLutX, lut, xl0, xih);
    node_lut(st->lnq, kLutQ, lut, ql0, qih);
    if (e == hipSuccess) e = hipMalloc(&st->d64, sizeof(double) * h64.size());
    if (e == hipSuccess) e = hipMalloc(&st->d32, sizeof(float) * h32.size());
    if (e == hipSuccess) e = hipMalloc(&st->dlut, sizeof(unsigned short) * lut.size());
    if (e == hipSuccess) e = hipMemcpy(st->d64, h64.data(), sizeof(double) * h64.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(st->d32, h32.data(), sizeof(float) * h32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(st->dlut, lut.data(), sizeof(unsigned short) * lut.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(st->d64); (void)hipFree(st->d32); (void)hipFree(st->dlut);
        delete st;
        return pfail(ctx, MI_ERR_HIP, std::string("PDF grid upload: ") + hipGetErrorString(e));
    }
    fill_view(st->g64, *st);
    st->g64.xk = st->d64; st->g64.lnx = st->d64 + nx; st->g64.qk = st->d64 + 2 * nx;
    st->g64.lnq = st->d64 + 2 * nx + nq; st->g64.val = st->d64 + 2 * nx + 2 * nq;
    fill_view(st->g32, *st);
    st->g32.val = st->d32;
    st->g32.ixinv = st->d32 + nv;
    st->g32.iqinv = st->g32.ixinv + kGridNF * nx;
    st->g32.xk = st->g32.iqinv + kGridNF * nq;
    st->g32.lnx = st->g32.xk + nx;
    st->g32.qk = st->g32.lnx + nx;
    st->g32.lnq = st->g32.qk + nq;
    st->g32.xlut = st->dlut;
    st->g32.qlut = st->dlut + kLutX;
    st->g32.xl0 = xl0; st->g32.xih = xih; st->g32.nbx = kLutX;
    st->g32.ql0 = ql0; st->g32.qih = qih; st->g32.nbq = kLutQ;
    fill_view(st->host64, *st);
    st->host64.xk = st->x.data(); st->host64.lnx = st->lnx.data(); st->host64.qk = st->qk.data();
    st->host64.lnq = st->lnq.data(); st->host64.val = st->val.data();
    PdfState *&dst = slot == 0 ? ctx->pdf : ctx->pdf_var[slot];
    release_state(dst);
    dst = st;
    return MI_OK;
}

int mi_pdf_xf(mi_ctx *ctx, uint64_t n, const double *x_dev, const double *Q2_dev, double *xf_dev)
{
    if (!ctx) return MI_ERR_ARG;
    if (!ctx->pdf) return pfail(ctx, MI_ERR_STATE, "no PDF grid loaded (mi_pdf_load)");
    if (n == 0) return MI_OK;
    if (!x_dev || !Q2_dev || !xf_dev) return pfail(ctx, MI_ERR_ARG, "NULL array");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_pdf_xf, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->pdf->g64, n,
                           x_dev, Q2_dev, xf_dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return pfail(ctx, MI_ERR_HIP, std::string("mi_pdf_xf: ") + hipGetErrorString(e));
    return MI_OK;
}

}  // extern "C"
