// mi_internal.h -- host-side context and kernel launchers shared by the library's translation
// units.  Not part of the C ABI (include/mcint.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/mcint.h"
#include "mi_device.h"

namespace mi {

// Block partial record of the integration kernels (FP64), kRecN = 1 + kMaxProc slots per
// quantity (0 = total, 1 + i = process slot i: codes 111..116, 121..124):
//   [kRS ..]  sum w        [kRQ ..] sum w^2        [kRM ..] max w
//   [kRNz]    points with total weight > 0
//   [kRAcc]   accepted candidates (unweighting)
constexpr int kRecN = kMaxProc + 1;
constexpr int kRS = 0;
constexpr int kRQ = kRecN;
constexpr int kRM = 2 * kRecN;
constexpr int kRNz = 3 * kRecN;
constexpr int kRAcc = kRNz + 1;
constexpr int kRec = 36;
static_assert(kRAcc < kRec, "record layout");
__host__ __device__ constexpr bool rec_is_max(int i) { return i >= kRM && i < kRM + kRecN; }
// Does mask select a heavy-quark code (121..124)?
__host__ __device__ constexpr bool has_heavy(int mask) { return (mask & (kFamMask[1] | kFamMask[2])) != 0; }
// Pythia process code of slot i.
__host__ __device__ constexpr int proc_code(int i) { return i < kLightProc ? 111 + i : 121 + (i - kLightProc); }
constexpr int kBlock = 256;
constexpr uint64_t kUnwChunk = 1ULL << 27;   // unweighting candidates per chunk (stored weights)
constexpr int kMaxBins = 256;
// Bytes of one wave's LDS area for the ordered VEGAS histogram over dims dimensions: the
// dims x nbins FP64 histogram, then one u64 lane mask per bin.
__host__ __device__ constexpr size_t wave_hist_bytes(int dims, int nbins)
{
    return (size_t)(dims + 1) * nbins * 8;
}

struct Grid {
    int ndim = 0, nbins = 0;
    double *edges = nullptr;   // ndim x (nbins + 1), FP64 (device)
    float *tab32 = nullptr;    // ndim x nbins x {lo, width, 1 - lo} (device)
    double *tab64 = nullptr;   // same in FP64 (device)
};

struct FlowState;              // mi_flow.hip
struct PdfState;               // mi_pdf.hip

}  // namespace mi

struct mi_ctx {
    int device = 0;
    int num_cu = 0;
    int lds_max = 0;             // LDS bytes a block may use (hipDeviceAttributeMaxSharedMemoryPerBlock)
    hipStream_t stream = nullptr;
    std::string err;
    // scratch, grown on demand (never shrunk)
    double *partials = nullptr;  size_t partials_cap = 0;   // doubles
    double *hist = nullptr;      size_t hist_cap = 0;       // doubles
    double *red = nullptr;       size_t red_cap = 0;        // doubles (reduced records)
    uint64_t *counts = nullptr;  size_t counts_cap = 0;     // per-block counts / offsets
    uint64_t *masks = nullptr;   size_t masks_cap = 0;      // unweighting accept ballots
    unsigned char *wbuf = nullptr; size_t wbuf_cap = 0;     // unweighting per-candidate weights (bytes)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    double *scratch = nullptr;   // device address of mi_integrate_noalloc's static scratch on this device
    mi::Grid grid;
    mi::FlowState *flow = nullptr;
    mi::PdfState *pdf = nullptr;                          // slot 0: the grid of pdf_source = MI_PDF_GRID
    mi::PdfState *pdf_var[MI_PDF_SLOTS] = {};            // slots 1..: variation grids (mi_reweight)
};

namespace mi {

// ------------------------------------------------------------------ constants
void fill_had_consts_f32(const mi_physics &p, HadConst<float> &k);
void fill_had_consts_f64(const mi_physics &p, HadConst<double> &k);

// ------------------------------------------------------------------ PDF grid (mi_pdf.hip)
// Views of the loaded grid for p (active only when p.pdf_source == MI_PDF_GRID); MI_ERR_STATE
// when a grid is asked for and none is loaded.
int pdf_views(mi_ctx *ctx, const mi_physics &p, GridPdf<float> *g32, GridPdf<double> *g64);
// The rows and columns of the FP32 grid that an integration of p reaches (x >= 4 pTmin^2 / S,
// q from fact_fac pTmin to fact_fac sqrt(S) / 2) as an LDS window of the integration kernel at
// byte offset off (sets g's woff, wq0, wq1, wx0, wstride); false, g unchanged, when p uses no grid
// or the window would end beyond max_bytes.  *bytes: the window's size.
bool pdf_window(const mi_ctx *ctx, const mi_physics &p, size_t off, size_t max_bytes, GridPdf<float> *g,
                size_t *bytes);
// Host (FP64) view for host-side evaluation; false when p does not use a grid.
bool pdf_host_view(const mi_ctx *ctx, const mi_physics &p, GridPdf<double> *g);
void pdf_release(mi_ctx *ctx);
// FP64 device view of PDF slot s (0 = the grid of mi_pdf_load); MI_ERR_STATE if none is loaded.
int pdf_slot_view(mi_ctx *ctx, int slot, GridPdf<double> *g64);
constexpr int kPdfSlots = MI_PDF_SLOTS;

// ------------------------------------------------------------------ launch geometry
struct Geom { uint32_t blocks; uint32_t ppt; };
// Points per thread >= min_ppt, blocks <= max_blocks.
Geom geom_for(uint64_t n, uint32_t min_ppt, uint32_t max_blocks, uint32_t nt = kBlock);
// Resident blocks per CU of the integration kernels (hipOccupancyMaxActiveBlocksPerMultiprocessor).
int occ_integrate_had(int prec, bool vegas, bool hist, int nbins, bool grid, bool heavy);
int occ_integrate_c1(bool vegas, bool hist, int nbins);

// ------------------------------------------------------------------ launchers (mi_kernels.hip)
hipError_t launch_philox_fill(hipStream_t s, uint64_t key, uint64_t ctr0, uint32_t stream_id,
                              uint64_t n, uint32_t *out);
hipError_t launch_sigma_hat(hipStream_t s, int code, uint64_t n, const double *sH, const double *tH,
                            const double *uH, const double *alpS, const int32_t *id1,
                            const int32_t *id2, int nQuarkNew, double kfactor, double *out);

// Integration of the hadronic integrand: writes blocks x kRec partials; if hist != nullptr
// also blocks x (ndim * nbins) per-bin sums of w^2 (VEGAS).
struct HadLaunch {
    int prec;            // MI_PREC_FP32 / FP64
    bool vegas;          // map through grid
    int mask;
    uint64_t key, ctr0, n;
    Geom g;
    HadConst<float> kf;
    HadConst<double> kd;
    const Grid *grid;
    double *partials;
    double *hist;
    GridPdf<float> gf;     // PDF grid (active = 0: toy PDFs)
    GridPdf<double> gd;
    bool win = false;      // FP32 grid with its LDS window (gf.woff): 1024-thread blocks
    size_t win_smem = 0;   // dynamic LDS bytes of the window launch (tables + window)
    bool hw = false;       // FP32 with the hardware transcendentals (run precision MI_PREC_FP32_HW)
};
constexpr int kWinBlock = 512;                 // threads per block of the LDS-window launch
hipError_t launch_integrate_had(hipStream_t s, const HadLaunch &L);

struct C1Launch {
    bool vegas;
    uint64_t key, ctr0, n;
    Geom g;
    double pref, cMax;
    const Grid *grid;
    double *partials;
    double *hist;
};
hipError_t launch_integrate_c1(hipStream_t s, const C1Launch &L);

// Per-point dump (parity hook).
hipError_t launch_points_had(hipStream_t s, const HadLaunch &L, double *w, double *kin);
hipError_t launch_points_c1(hipStream_t s, const C1Launch &L, double sHat, double *w, double *kin);

// Deterministic reductions.
hipError_t launch_reduce_records(hipStream_t s, const double *partials, uint32_t nblk, double *out);

// VEGAS refinement on the device (one block).
// One VEGAS iteration's reductions and refinement in two launches: the nblk block records into
// rec_out (kRec doubles) and the nblk x (ndim * nbins) histogram rows into hsum, then the grid
// edges and tables refined from hsum.
hipError_t launch_vegas_step(hipStream_t s, const double *partials, uint32_t nblk, double *rec_out,
                             const double *hist_parts, Grid &g, double *hsum, double alpha);
hipError_t launch_vegas_tables(hipStream_t s, Grid &g);

// Unweighting (mi_unweight.hip).
// One chunk of candidates [ctr0, ctr0 + n) (h or c carries them and the chunk's geometry).
struct UnwLaunch {
    HadLaunch h;          // hadronic candidates (h.partials receives the estimate partials)
    bool partonic;        // partonic gg->gg (C1 integrand) instead
    C1Launch c;
    double sHat;          // partonic
    bool compat;          // SplitMix64 reference-compat stream (partonic only)
    double wmax;
    void *w;              // per-candidate total weights (T), chunk-local index; eval may take null
    uint64_t *masks;      // blocks x ppt x 4 wave ballots
    uint64_t *counts;     // blocks + 1: accepted per block, then scanned (counts[blocks] = total)
    uint64_t base;        // rank of the chunk's first accepted event
    uint32_t mat_blocks;  // grid of the materialize launch (grid-stride over the accepted events)
    const mi_events *out;
    bool win = false;     // FP32 grid: the eval kernel keeps the window of wgrid in LDS
    size_t win_smem = 0;  // its dynamic LDS bytes (tables + window)
    GridPdf<float> wgrid; // the window (woff, wq0, wq1, wx0, wstride) for the eval kernel only
};
hipError_t launch_unweight_eval(hipStream_t s, const UnwLaunch &L);
int occ_unweight_eval(const UnwLaunch &L);   // resident blocks per CU of the eval kernel
hipError_t launch_unweight_flags(hipStream_t s, const UnwLaunch &L);
hipError_t launch_scan_counts(hipStream_t s, uint64_t *counts, uint32_t nblk);
hipError_t launch_unweight_compact(hipStream_t s, const UnwLaunch &L);
hipError_t launch_unweight_materialize(hipStream_t s, const UnwLaunch &L);

// Scale and PDF variations of events (mi_kernels.hip): per variation xi_R^2, xi_F^2 and a PDF slot.
struct RwVar { double xiR2, xiF2; int slot, pad; };
struct GridSet64 { GridPdf<double> g[MI_PDF_SLOTS]; };
hipError_t launch_reweight_vars(hipStream_t s, const HadConst<double> &k, const GridSet64 &gs, bool grid, uint64_t n,
                                const double *sH, const double *tH, const double *x1, const double *x2,
                                const int32_t *pid, const double *w, int nvar, const RwVar *vars_dev,
                                double xiR2_0, double xiF2_0, double *out);
// Device event record (mi_events.hip).
struct EvConst {
    double eBeam;                        // per beam (partonic: sqrt(sHat) / 2)
    double renorm_fac, nQuarkNew;
    double Ag, ag, bg, Auv, Adv, av, bv, As, as, bs;   // toy xf = A x^a (1 - x)^b
    double mQ[2];
    int nQ, partonic, grid, max_part;
};
// Constants of the event build for p (validates p; the grid's device view when p uses one).
int event_consts(mi_ctx *ctx, const mi_physics &p, int max_part, EvConst *c, GridPdf<double> *g);
// Builds events first .. first + n - 1 of ev into record slots 0 .. n - 1 of out.
hipError_t launch_build_events(hipStream_t s, const EvConst &c, const GridPdf<double> &g, uint64_t n, uint64_t key,
                               const mi_events &ev, uint64_t first, const mi_particles &out);

// Renormalisation-scale reweighting of events (mi_kernels.hip).
hipError_t launch_reweight(hipStream_t s, const HadConst<double> &k, uint64_t n, const double *sH, const double *tH,
                           const int32_t *pid, const double *w, int nvar, const double *fac, double *out);

// Variations of events regenerated from their counters (mi_kernels.hip): per variation the run's
// FP64 constants at the varied scales and the FP64 view of its PDF slot (active = 0: toy PDFs).
// vars_dev holds nvar + 1 entries, [0] = the run's own.
struct RwRegen { HadConst<double> k; GridPdf<double> g; };
hipError_t launch_reweight_regen(hipStream_t s, const RwRegen *vars_dev, int nvar, int mask, bool vegas,
                                 uint64_t key, uint64_t n, const uint64_t *idx, const int32_t *pid, const double *w,
                                 const double *gtab, int nbins, double *out);

}  // namespace mi
