// librln_estep: projector / backprojector objects and the batched E-step engine behind the C ABI of
// include/rln_estep.h. Host C++ driving the HIP kernels of kernels.h on one device and one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"
#include "rln_common.h"
#include "sampling.h"

namespace rln {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const char *last_error() { return g_last_error.c_str(); }

}  // namespace rln

#include "host_common.h"

namespace rln {
// Packs a plain [z][y][x] float2 volume already on the device into the projector's corner layout.
int pack_device_into_projector(rln_projector *p, const float2 *dev_plain, hipStream_t st) {
    const rln::MdlDims &d = p->dims;
    const rlnk::VolGeom g{d.xdim, d.ydim, d.zdim, d.inity, d.initz};
    const size_t n = p->bases();
    hipLaunchKernelGGL(rlnk::k_pack_volume, dim3((unsigned)((n + rlnk::kBlock - 1) / rlnk::kBlock)),
                       dim3(rlnk::kBlock), 0, st, dev_plain, g, p->data.as<float4>());
    HIP_TRY(hipGetLastError());
    p->version++;
    return RLN_OK;
}

// Sums the backprojector's corner slots into plain [z][y][x] data (re, im) and weight volumes on the device.
int fold_backprojector_device(rln_backprojector *b, float *re, float *im, float *w, hipStream_t st) {
    const rlnk::VolGeom g{b->dims.xdim, b->dims.ydim, b->dims.zdim, b->dims.inity, b->dims.initz};
    hipLaunchKernelGGL(rlnk::k_fold_slots, dim3(rlnk::fold_blocks(g)), dim3(rlnk::kBlock), 0, st,
                       b->slots.as<const float>(), g, 0, re, im, w);
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

// Uploads a plain [z][y][x] float2 volume and packs it into the projector's corner layout.
static int pack_into_projector(rln_projector *p, const float2 *host_plain) {
    const rln::MdlDims &d = p->dims;
    DevBuf plain;
    RLN_TRY(plain.ensure(sizeof(float2) * d.voxels()));
    HIP_TRY(hipMemcpy(plain.ptr, host_plain, sizeof(float2) * d.voxels(), hipMemcpyHostToDevice));
    const rlnk::VolGeom g{d.xdim, d.ydim, d.zdim, d.inity, d.initz};
    const size_t n = p->bases();
    hipLaunchKernelGGL(rlnk::k_pack_volume, dim3((unsigned)((n + rlnk::kBlock - 1) / rlnk::kBlock)),
                       dim3(rlnk::kBlock), 0, 0, plain.as<const float2>(), g, p->data.as<float4>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return RLN_OK;
}
}  // namespace rln

namespace rln {

struct ChunkHost;

// Allocator whose value-less construct leaves trivial elements uninitialised: resize() of a host list that is
// written in full afterwards costs no zero fill.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U> &) {}
    template <class U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostList = std::vector<T, UninitAlloc<T>>;

// Pixel lists of the banded kernels: the pixels `pix` (pixel order) of an s x xs half-plane ranked by radius (ties
// in pixel order), cut into nb equal parts for every band choice, each band in pixel order, stored as
// (p, x | row << 16): kNumBandChoices lists of pix.size() entries.
static std::vector<int2> banded_lists(const std::vector<int> &pix, int s, int xs);

// The production fine diff2 kernel of a work-item shape (index into rlnk::kFineShapes).
inline const void *fine_kernel_ptr(int shape) {
    switch (shape) {
        case 1: return reinterpret_cast<const void *>(&rlnk::k_fine_diff2<0, 1, 2>);
        case 2: return reinterpret_cast<const void *>(&rlnk::k_fine_diff2<0, 2, 1>);
        case 3: return reinterpret_cast<const void *>(&rlnk::k_fine_diff2<0, 1, 1>);
        default: return reinterpret_cast<const void *>(&rlnk::k_fine_diff2<0, 2, 2>);
    }
}

// Band counts the fine pass can split its sampled pixels into (radial shells of equal pixel count; the tuner
// picks one per engine).
constexpr int kBandChoices[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumBandChoices = (int)(sizeof(kBandChoices) / sizeof(kBandChoices[0]));

// Orders of the fine pass's work units: jobs in particle order or sorted by class and coarse orientation
// (k_fine_items), units band-major or item-major, in one queue or in per-XCD queues (k_fine_diff2).
struct FineOrder {
    int sorted, item_major, xcd_queues;
};
constexpr FineOrder kFineOrders[] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 0, 1}};
constexpr int kNumFineOrders = (int)(sizeof(kFineOrders) / sizeof(kFineOrders[0]));

struct Engine;
static int wsum_band_choice(const Engine &e);

static std::vector<int2> banded_lists(const std::vector<int> &pix, int s, int xs) {
    const int ns = (int)pix.size();
    auto r2 = [&](int p) {
        int x, y;
        rlnk::pixel_xy(p, s, xs, x, y);
        return x * x + y * y;
    };
    std::vector<int> byr(pix);
    std::stable_sort(byr.begin(), byr.end(), [&](int a, int b) { return r2(a) < r2(b); });
    std::vector<int> band_p(ns);
    std::vector<int2> lists((size_t)kNumBandChoices * std::max(ns, 1));
    for (int c = 0; c < kNumBandChoices; ++c) {
        const int nb = kBandChoices[c];
        for (int band = 0; band < nb; ++band) {
            const int i0 = band * ns / nb, i1 = (band + 1) * ns / nb;
            std::copy(byr.begin() + i0, byr.begin() + i1, band_p.begin() + i0);
            std::sort(band_p.begin() + i0, band_p.begin() + i1);
        }
        // the kernels read the pixel's coordinates instead of dividing by the row length
        int2 *dst = lists.data() + (size_t)c * ns;
        for (int k = 0; k < ns; ++k) {
            const int p = band_p[k], row = p / xs;
            dst[k] = make_int2(p, (p - row * xs) | (row << 16));
        }
    }
    return lists;
}

// Stage timers of one run (HIP events on the engine's stream).
enum Stage { kStPrep, kStCoarse, kStCoarseW, kStJobs, kStFine, kStFineW, kStWsum, kStCount };
// fine stage graphs: plan and work items; fine diff2; fine weights; weighted sums and backprojection
constexpr int kFineGraphs = 4;

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    rln_estep_config cfg{};
    PoseGrid grid;
    int NT = 0, n_or = 0;
    int sc = 0, sf = 0;
    rlnk::ImgGeom igc{}, igf{};
    rlnk::VolGeom vg{};
    int nshell = 0;
    // model
    int K = 0;
    std::vector<rln_projector *> proj;
    std::vector<rln_backprojector *> bproj;
    std::vector<unsigned long long> proj_version;
    std::vector<double> pdf_class, pdf_direction, sigma2_noise, data_vs_prior;
    double pdf_dir_mean = 1.;            // mean of pdf_direction over classes and directions
    std::vector<int> set_off, set_og;    // explicit orientation lists for the next run (empty: none)
    bool model_set = false;
    bool coarse_valid = false;
    // constant tables
    DevBuf d_frames_c, d_frames_f, d_proj_c;
    DevBuf d_pf, d_zero_row;                        // split-BF16 coarse projections (FP32 complex), a zero row
    DevBuf d_ehi, d_elo;                            // split-BF16 translation phases [NT][Kp/2]
    DevBuf d_wc, d_ctf2, d_ynorm;                   // split-BF16 particle rows per chunk (group layout)
    DevBuf d_grp, d_exg;                            // pixel groups (row, x0); x phases with rows padded to xs4
    int Kp = 0;                                     // contraction length: 8 * groups, multiple of 32
    int G = 0, xs4 = 0;                             // pixel groups with weight; coarse row length padded to 4
    DevBuf d_q_c, d_q_f, d_minv_w, d_mresol, d_xa_shell;
    DevBuf d_ex_c, d_ey_c, d_ex_f, d_ey_f, d_tr_ix, d_tr_iy, d_over_x, d_over_y;
    DevBuf d_vols, d_bp_ptrs;
    std::vector<double> over_x, over_y;  // [NT*4] pixels
    // coarse translations on their lattice: distinct x and y values (Angstrom) and each translation's indices
    std::vector<double> tr_ux, tr_uy;
    std::vector<int> tr_cix, tr_ciy;
    std::vector<int> mresol_f;
    double logsigma2 = 0.;
    // per-chunk buffers
    DevBuf d_fimg, d_fimgn, d_fctf;
    DevBuf d_meta_f, d_meta_d, d_meta_ll;
    DevBuf d_or_off, d_or_cnt, d_or_proj, d_or_prior, d_or_frame, d_or_class, d_or_dir, d_tr_prior;
    DevBuf d_oc_off;  // per-particle offsets of per-orientation state
    DevBuf d_ctfq, d_yq, d_tiles;
    DevBuf d_diff2, d_weight, d_mask, d_stats_c, d_stats_f;
    DevBuf d_npairs, d_njobs, d_nchunks, d_pair_off, d_job_off, d_chunk_off, d_pairs, d_jobs, d_chunk_job,
        d_chunk_first;
    DevBuf d_fdiff2, d_fweight, d_fine_n, d_nent, d_foff;
    DevBuf d_fbase;  // pose-independent part of each particle's fine diff2 (FP64)
    DevBuf d_shell_acc, d_part_acc, d_pdfc_acc, d_pdfd_acc;
    DevBuf d_shell_tot;  // [nshell] the chunk's sigma2_noise sums over its particles
    // fine pass planned on the device: capacities (elements) the buffers were sized for, the chunk capacity B,
    // the plan header, per-chunk image pointers, work items, the tuned persistent grids and the captured graph
    int cap_B = 0;
    long long need_pairs = 0;
    int need_jobs = 0, need_small = 0, need_wide = 0;
    DevBuf d_plan, d_cp, d_items_s, d_items_w;
    DevBuf d_wsegs, d_seg_first, d_wpart, d_hcnt, d_hsum, d_wcand, d_segbest, d_ocnt;  // segmented coarse weights
    PinnedBuf h_stage[2];  // a chunk's lists packed for upload, alternating between consecutive chunks
    std::unique_ptr<ChunkHost> chunk_host[2];  // host lists of consecutive chunks, kept to reuse their memory
    rlnk::ChunkPtrs cp_host{};
    bool tuned = false;
    int grid_fine = 1024, grid_wsum = 2048;  // persistent grids (fine: tuned; weighted sums: resident blocks)
    int band_choice = 0;                     // index into kBandChoices of the fine pass's band count (tuned)
    int fine_shape = 0;                      // index into rlnk::kFineShapes of the work-item shape (tuned)
    int fine_order = 0;                      // order of the fine units (kFineOrders; tuned)
    int order_fixed = -1;                    // fine order set by rln_estep_set_fine_order (-1: tuned)
    int shape_fixed = -1;                    // work-item shape set by rln_estep_set_fine_shape (-1: tuned)
    int wsum_order = 0;                 // unit order of the weighted sums: 0 band-major, 1 item-major, 2 per-XCD queues
    int wsum_bands = 0;                      // band count of the weighted sums (rln_estep_set_wsum_bands; 0: the fine pass's)
    DevBuf d_job_order;                      // jobs in the fine pass's order (k_fine_items)
    int resident_shape[rlnk::kNumFineShapes] = {};  // resident blocks of each shape's kernel on the device
    DevBuf d_band_pix;                       // sampled fine pixels, one banded list per band choice
    DevBuf d_wsum_pix;                       // pixels of the weighted sums, one banded list per band choice
    int n_wsum_pix = 0;
    std::vector<int> band_src;               // the sampled pixels the lists were built from (pixel order)
    int resident_fine = 0;                   // resident blocks of the default shape's k_fine_diff2
    int n_cu = 0;                            // compute units of the device
    hipGraph_t fine_graph[kFineGraphs] = {};
    hipGraphExec_t fine_exec[kFineGraphs] = {};
    std::vector<unsigned char> fine_key;  // kernel arguments the graph was captured with
    bool graph_off = false;               // capture unsupported: launch directly
    // intermediates kept for tests
    bool keep = false;
    int max_pool = 16384;  // particles per chunk at most (rln_estep_set_pool)
    struct Kept {
        long long n_coarse = 0;
        int n_dir = 0, n_psi = 0;
        std::vector<float> cdiff2;
        std::vector<double> cweight;
        std::vector<unsigned char> cmask;
        std::vector<long long> f_index;
        std::vector<double> fdiff2;
        std::vector<double> fweight;
    };
    std::vector<Kept> kept;
    // timings
    hipEvent_t ev[kStCount + 1] = {};
    double stage_ms[kStCount] = {};
    double host_ms[4] = {};  // host phases: chunk preparation, of it the first chunk's lists, epilogue, whole run
    long long counters[23] = {};
    long long pf_sampled = 0;  // fine pixels gathered per fine orientation
    std::vector<double> tune_ms;  // fine-pass tuning: time of each (band choice, grid) candidate, -1 untried
    bool have_timing = false;
    std::mutex mtx;
};

// Band choice of the weighted sums: the one set by rln_estep_set_wsum_bands, else the fine pass's.
static int wsum_band_choice(const Engine &e) {
    if (e.wsum_bands > 0)
        for (int c = 0; c < kNumBandChoices; ++c)
            if (kBandChoices[c] == e.wsum_bands) return c;
    return e.band_choice;
}

namespace {

// Mresol of an s x (s/2+1) image (src/ml_optimiser.cpp:5782-5812).
std::vector<int> mresol_of(int s) {
    const int xs = s / 2 + 1;
    std::vector<int> m((size_t)s * xs, -1);
    for (int i = 0; i < s; ++i) {
        const int ip = (i < xs) ? i : i - s;
        for (int j = 0; j < xs; ++j) {
            const int ires = ROUND(sqrt((double)(ip * ip + j * j)));
            if (ires < s / 2 + 1 && !(j == 0 && ip < 0)) m[(size_t)i * xs + j] = ires;
        }
    }
    return m;
}

rlnk::ImgGeom img_geom(int s, int r_max, double pad) {
    rlnk::ImgGeom g;
    g.s = s;
    g.xs = s / 2 + 1;
    g.P = s * g.xs;
    const int rmo = g.xs - 1;
    const long long rmr = (long long)(r_max * pad);
    const double pad2 = pad * pad;
    int r2 = rmo * rmo;
    while (r2 > 0 && pad2 * (double)r2 > (double)(rmr * rmr)) --r2;
    g.proj_r2_max = r2;
    const long long rr = ROUND(r_max * pad);
    int b2 = (int)std::min<long long>(rr * rr, 4LL * s * s);
    while (b2 > 0 && pad2 * (double)b2 > (double)(rr * rr)) --b2;
    g.bp_r2_max = b2;
    return g;
}

rlnk::Frame frame_of(double rot, double tilt, double psi, double pad) {
    double A[9];
    euler_angles2matrix(rot, tilt, psi, A);
    rlnk::Frame f;
    f.e0x = (float)(pad * A[0]);
    f.e0y = (float)(pad * A[1]);
    f.e0z = (float)(pad * A[2]);
    f.e1x = (float)(pad * A[3]);
    f.e1y = (float)(pad * A[4]);
    f.e1z = (float)(pad * A[5]);
    return f;
}

// exp(-2 pi i k shift / N) for k on the axis of an image (x: 0..xs-1; y rows: FFTW signed order).
void phase_rows(const std::vector<double> &shifts, int s, bool yaxis, double N, std::vector<float2> &out) {
    const int xs = s / 2 + 1;
    const int len = yaxis ? s : xs;
    out.resize(shifts.size() * (size_t)len);
    for (size_t j = 0; j < shifts.size(); ++j)
        for (int i = 0; i < len; ++i) {
            const int k = yaxis ? ((i < xs) ? i : i - s) : i;
            // reduce the phase in turns before the trigonometry
            double turns = -(double)k * shifts[j] / N;
            turns -= std::nearbyint(turns);
            const double a = 2. * kPi * turns;
            out[j * len + i] = make_float2((float)cos(a), (float)sin(a));
        }
}

template <class T>
int upload(DevBuf &buf, const std::vector<T> &v, hipStream_t st) {
    RLN_TRY(buf.ensure(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return RLN_OK;
}

int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) RLN_FAIL(RLN_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return RLN_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Engine setup
// ------------------------------------------------------------------------------------------------
static int engine_init(Engine &e, const rln_estep_config &cfg) {
    e.cfg = cfg;
    if (cfg.nr_classes <= 0 || cfg.ori_size <= 0 || (cfg.ori_size & 1) || cfg.current_size <= 0 ||
        (cfg.current_size & 1) || cfg.coarse_size <= 0 || (cfg.coarse_size & 1) || cfg.coarse_size > cfg.current_size ||
        cfg.current_size > cfg.ori_size || !(cfg.padding_factor >= 1.))
        RLN_FAIL(RLN_ERR_ARG, "invalid image sizes (ori %d, current %d, coarse %d) or padding %g", cfg.ori_size,
                 cfg.current_size, cfg.coarse_size, cfg.padding_factor);
    if (cfg.adaptive_oversampling != 1) RLN_FAIL(RLN_ERR_ARG, "only adaptive_oversampling = 1 is supported");
    if (cfg.precision != RLN_PREC_PLAIN_FP32 && cfg.precision != RLN_PREC_SPLIT_BF16)
        RLN_FAIL(RLN_ERR_ARG, "unsupported precision mode %d", cfg.precision);
    if (!(cfg.sigma2_fudge > 0.) || !(cfg.adaptive_fraction > 0. && cfg.adaptive_fraction <= 1.))
        RLN_FAIL(RLN_ERR_ARG, "invalid sigma2_fudge or adaptive_fraction");
    if (cfg.ori_size / 2 + 1 > 512) RLN_FAIL(RLN_ERR_ARG, "ori_size %d above the supported 1022", cfg.ori_size);
    if (!e.grid.build(cfg)) RLN_FAIL(RLN_ERR_ARG, "invalid sampling options");
    e.NT = e.grid.n_trans;
    if (4 * e.NT > rlnk::kMaxSigT)
        RLN_FAIL(RLN_ERR_ARG, "%d translations: at most %d are supported", e.NT, rlnk::kMaxSigT / 4);
    e.n_or = e.grid.n_dir * e.grid.n_psi;
    e.sc = cfg.coarse_size;
    e.sf = cfg.current_size;
    const int r_max = std::min(cfg.current_size / 2, cfg.ori_size / 2);
    e.igc = img_geom(e.sc, r_max, cfg.padding_factor);
    e.igf = img_geom(e.sf, r_max, cfg.padding_factor);
    e.nshell = cfg.ori_size / 2 + 1;
    HIP_TRY(hipSetDevice(e.device));
    HIP_TRY(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    for (int i = 0; i <= kStCount; ++i) HIP_TRY(hipEventCreate(&e.ev[i]));
    // persistent grids sized to the blocks the device holds at once (dynamic work fetch inside)
    {
        int cus = 0, nb_f = 0, nb_w = 0;
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e.device));
        for (int sh = 0; sh < rlnk::kNumFineShapes; ++sh) {
            int nb = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fine_kernel_ptr(sh), rlnk::kBlock, 0));
            e.resident_shape[sh] = nb;
        }
        nb_f = e.resident_shape[0];
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_w, rlnk::k_wsum_bp, rlnk::kBlock, 0));
        e.resident_fine = cus * nb_f;
        e.n_cu = cus;
        if (cus * nb_w > 0) e.grid_wsum = cus * nb_w;
    }
    // orientation frames of the coarse grid and of its order-1 oversampling
    std::vector<rlnk::Frame> fc(e.n_or), ff((size_t)e.n_or * 8);
    double r8[8], t8[8], p8[8];
    for (int d = 0; d < e.grid.n_dir; ++d)
        for (int p = 0; p < e.grid.n_psi; ++p) {
            const int o = d * e.grid.n_psi + p;
            fc[o] = frame_of(e.grid.rot[d], e.grid.tilt[d], e.grid.psi[p], cfg.padding_factor);
            e.grid.oversampled_orientations(d, p, r8, t8, p8);
            for (int r = 0; r < 8; ++r) ff[(size_t)o * 8 + r] = frame_of(r8[r], t8[r], p8[r], cfg.padding_factor);
        }
    RLN_TRY(upload(e.d_frames_c, fc, e.stream));
    RLN_TRY(upload(e.d_frames_f, ff, e.stream));
    // translations: coarse phase rows per translation, fine rows per distinct sub-shift
    std::vector<double> cx(e.NT), cy(e.NT);
    for (int t = 0; t < e.NT; ++t) {
        cx[t] = e.grid.tx[t] / cfg.pixel_size;
        cy[t] = e.grid.ty[t] / cfg.pixel_size;
    }
    {
        std::map<double, int> mx, my;
        for (int t = 0; t < e.NT; ++t) {
            mx.emplace(e.grid.tx[t], 0);
            my.emplace(e.grid.ty[t], 0);
        }
        e.tr_ux.clear();
        e.tr_uy.clear();
        for (auto &kv : mx) {
            kv.second = (int)e.tr_ux.size();
            e.tr_ux.push_back(kv.first);
        }
        for (auto &kv : my) {
            kv.second = (int)e.tr_uy.size();
            e.tr_uy.push_back(kv.first);
        }
        e.tr_cix.resize(e.NT);
        e.tr_ciy.resize(e.NT);
        for (int t = 0; t < e.NT; ++t) {
            e.tr_cix[t] = mx[e.grid.tx[t]];
            e.tr_ciy[t] = my[e.grid.ty[t]];
        }
    }
    std::vector<float2> rows;
    phase_rows(cx, e.sc, false, cfg.ori_size, rows);
    RLN_TRY(upload(e.d_ex_c, rows, e.stream));
    phase_rows(cy, e.sc, true, cfg.ori_size, rows);
    RLN_TRY(upload(e.d_ey_c, rows, e.stream));
    e.over_x.resize((size_t)e.NT * 4);
    e.over_y.resize((size_t)e.NT * 4);
    for (int t = 0; t < e.NT; ++t) e.grid.oversampled_translations(t, &e.over_x[t * 4], &e.over_y[t * 4]);
    std::map<double, int> ux, uy;
    for (double v : e.over_x) ux.emplace(v, 0);
    for (double v : e.over_y) uy.emplace(v, 0);
    std::vector<double> vx, vy;
    for (auto &kv : ux) {
        kv.second = (int)vx.size();
        vx.push_back(kv.first);
    }
    for (auto &kv : uy) {
        kv.second = (int)vy.size();
        vy.push_back(kv.first);
    }
    std::vector<int> tix(e.NT * 4), tiy(e.NT * 4);
    for (int i = 0; i < e.NT * 4; ++i) {
        tix[i] = ux[e.over_x[i]];
        tiy[i] = uy[e.over_y[i]];
    }
    phase_rows(vx, e.sf, false, cfg.ori_size, rows);
    RLN_TRY(upload(e.d_ex_f, rows, e.stream));
    phase_rows(vy, e.sf, true, cfg.ori_size, rows);
    RLN_TRY(upload(e.d_ey_f, rows, e.stream));
    RLN_TRY(upload(e.d_tr_ix, tix, e.stream));
    RLN_TRY(upload(e.d_tr_iy, tiy, e.stream));
    RLN_TRY(upload(e.d_over_x, e.over_x, e.stream));
    RLN_TRY(upload(e.d_over_y, e.over_y, e.stream));
    e.mresol_f = mresol_of(e.sf);
    RLN_TRY(upload(e.d_mresol, e.mresol_f, e.stream));
    {
        // pixels of the weighted sums: a resolution shell (noise spectrum) or inside the backprojection radius and
        // not on the duplicated x = 0, y < 0 half column (backproject2Dto3D, src/backprojector.cpp:134-143)
        std::vector<int> wp;
        for (int n = 0; n < e.igf.P; ++n) {
            int x, y;
            rlnk::pixel_xy(n, e.igf.s, e.igf.xs, x, y);
            const bool bp = x * x + y * y <= e.igf.bp_r2_max && !(n / e.igf.xs >= e.igf.xs && x == 0);
            if (e.mresol_f[n] > -1 || bp) wp.push_back(n);
        }
        e.n_wsum_pix = (int)wp.size();
        RLN_TRY(upload(e.d_wsum_pix, banded_lists(wp, e.igf.s, e.igf.xs), e.stream));
    }
    if (cfg.precision == RLN_PREC_SPLIT_BF16 && e.NT > 256)
        RLN_FAIL(RLN_ERR_ARG, "split-BF16 coarse pass supports at most 256 translations (%d)", e.NT);
    if (cfg.precision == RLN_PREC_SPLIT_BF16) {
        // x phases of the coarse translations with rows padded to a multiple of 4 (zeros past xs)
        e.xs4 = 4 * ((e.igc.xs + 3) / 4);
        phase_rows(cx, e.sc, false, cfg.ori_size, rows);
        std::vector<float2> rows4((size_t)e.NT * e.xs4, make_float2(0.f, 0.f));
        for (int t = 0; t < e.NT; ++t)
            for (int x = 0; x < e.igc.xs; ++x) rows4[(size_t)t * e.xs4 + x] = rows[(size_t)t * e.igc.xs + x];
        RLN_TRY(upload(e.d_exg, rows4, e.stream));
    }
    HIP_TRY(hipStreamSynchronize(e.stream));
    return RLN_OK;
}

// Per-model tables: Minvsigma2 at both sizes, volume pointers, backprojector pointers.
static int engine_model_tables(Engine &e) {
    const rln_estep_config &cfg = e.cfg;
    auto q_of = [&](int s, std::vector<float> &q, std::vector<float> *minv_w) {
        const std::vector<int> m = mresol_of(s);
        q.assign(m.size(), 0.f);
        if (minv_w) minv_w->assign(m.size(), 0.f);
        for (size_t n = 0; n < m.size(); ++n) {
            const int ires = m[n];
            double minv = 0.;
            if (ires > 0 && ires < e.nshell) minv = 1. / (cfg.sigma2_fudge * e.sigma2_noise[ires]);
            q[n] = (float)sqrt(0.5 * minv);
            if (minv_w) (*minv_w)[n] = (float)minv;
        }
        if (minv_w) (*minv_w)[0] = (float)(1. / (cfg.sigma2_fudge * e.sigma2_noise[0]));  // origin restored
    };
    std::vector<float> qc, qf, mw;
    q_of(e.sc, qc, nullptr);
    q_of(e.sf, qf, &mw);
    if (cfg.precision == RLN_PREC_SPLIT_BF16) {
        // pixel groups of the split pass: per row, x = 0 .. last pixel with non-zero Minvsigma2, in fours
        std::vector<int2> grp;
        for (int i = 0; i < e.igc.s; ++i) {
            int xmax = -1;
            for (int x = 0; x < e.igc.xs; ++x)
                if (qc[(size_t)i * e.igc.xs + x] != 0.f) xmax = x;
            for (int x0 = 0; x0 <= xmax; x0 += 4) grp.push_back(make_int2(i, x0));
        }
        e.G = (int)grp.size();
        const int ng = 4 * ((e.G + 3) / 4);
        grp.resize(std::max(ng, 4), make_int2(0, 0));
        e.Kp = 8 * (int)grp.size();
        RLN_TRY(upload(e.d_grp, grp, e.stream));
        std::vector<float2> z(e.Kp / 2, make_float2(0.f, 0.f));
        RLN_TRY(upload(e.d_zero_row, z, e.stream));
        const long long ne = (long long)e.NT * (e.Kp / 2);
        RLN_TRY(e.d_ehi.ensure(sizeof(unsigned int) * (size_t)ne));
        RLN_TRY(e.d_elo.ensure(sizeof(unsigned int) * (size_t)ne));
        hipLaunchKernelGGL(rlnk::k_phase_split, dim3((unsigned)((ne + rlnk::kBlock - 1) / rlnk::kBlock)),
                           dim3(rlnk::kBlock), 0, e.stream, e.d_exg.as<const float2>(), e.xs4,
                           e.d_ey_c.as<const float2>(), e.igc.s, e.d_grp.as<const int2>(), e.Kp / 2, e.NT,
                           e.d_ehi.as<unsigned int>(), e.d_elo.as<unsigned int>());
        HIP_TRY(hipGetLastError());
    }
    RLN_TRY(upload(e.d_q_c, qc, e.stream));
    RLN_TRY(upload(e.d_q_f, qf, e.stream));
    // fine pixels gathered per fine orientation: non-zero Minvsigma2 inside the projection radius
    std::vector<int> samp;
    for (int n = 0; n < e.igf.P; ++n) {
        int x, y;
        rlnk::pixel_xy(n, e.igf.s, e.igf.xs, x, y);
        if (qf[n] != 0.f && x * x + y * y <= e.igf.proj_r2_max) samp.push_back(n);
    }
    e.pf_sampled = (long long)samp.size();
    if (samp != e.band_src) {
        RLN_TRY(upload(e.d_band_pix, banded_lists(samp, e.igf.s, e.igf.xs), e.stream));
        e.band_src = samp;
    }
    RLN_TRY(upload(e.d_minv_w, mw, e.stream));
    // logsigma2 of the dLL (src/ml_optimiser.cpp:9024-9033)
    e.logsigma2 = 0.;
    for (int ires : e.mresol_f)
        if (ires > 0 && ires < e.nshell) e.logsigma2 += log(2. * kPi * e.sigma2_noise[ires]);
    std::vector<const float4 *> vp(e.K);
    std::vector<float *> bp(e.K);
    for (int k = 0; k < e.K; ++k) {
        vp[k] = e.proj[k]->data.as<const float4>();
        bp[k] = e.bproj[k]->slots.as<float>();
    }
    RLN_TRY(upload(e.d_vols, vp, e.stream));
    RLN_TRY(upload(e.d_bp_ptrs, bp, e.stream));
    if (cfg.do_scale_correction && !e.data_vs_prior.empty()) {
        std::vector<unsigned char> xa((size_t)e.K * e.nshell);
        for (size_t i = 0; i < xa.size(); ++i) xa[i] = e.data_vs_prior[i] > 3. ? 1 : 0;
        RLN_TRY(upload(e.d_xa_shell, xa, e.stream));
    }
    RLN_TRY(e.d_pdfd_acc.ensure(sizeof(double) * (size_t)e.K * e.grid.n_dir));
    return RLN_OK;
}

static int engine_coarse_projections(Engine &e) {
    bool stale = !e.coarse_valid;
    for (int k = 0; k < e.K; ++k)
        if (e.proj_version[k] != e.proj[k]->version) stale = true;
    if (!stale) return RLN_OK;
    if (e.cfg.precision == RLN_PREC_SPLIT_BF16) {
        const size_t rows = (size_t)e.K * e.n_or;
        RLN_TRY(e.d_pf.ensure(rows * (e.Kp / 2) * sizeof(float2)));
        dim3 grid(e.K * e.n_or, (e.Kp / 2 + rlnk::kBlock - 1) / rlnk::kBlock);
        hipLaunchKernelGGL(rlnk::k_project_grid_split, grid, dim3(rlnk::kBlock), 0, e.stream,
                           e.d_vols.as<const float4 *const>(), e.vg, e.d_frames_c.as<const rlnk::Frame>(), e.n_or,
                           e.igc, e.d_grp.as<const int2>(), e.G, e.Kp, e.d_pf.as<float2>());
        RLN_TRY(launch_check("k_project_grid_split"));
    } else {
        const size_t n = (size_t)e.K * e.n_or * e.igc.P;
        RLN_TRY(e.d_proj_c.ensure(n * sizeof(float2)));
        dim3 grid(e.K * e.n_or, (e.igc.P + rlnk::kBlock - 1) / rlnk::kBlock);
        hipLaunchKernelGGL(rlnk::k_project_grid, grid, dim3(rlnk::kBlock), 0, e.stream,
                           e.d_vols.as<const float4 *const>(), e.vg, e.d_frames_c.as<const rlnk::Frame>(), e.n_or,
                           e.igc, e.d_proj_c.as<float2>());
        RLN_TRY(launch_check("k_project_grid"));
    }
    for (int k = 0; k < e.K; ++k) e.proj_version[k] = e.proj[k]->version;
    e.coarse_valid = true;
    e.counters[8] += (long long)e.K * e.n_or;
    return RLN_OK;
}

// ------------------------------------------------------------------------------------------------
// One chunk of particles
// ------------------------------------------------------------------------------------------------
struct ChunkHost {
    int B = 0;
    std::vector<int> or_off, or_cnt;
    HostList<int> or_proj, or_frame, or_class, or_dir;  // per orientation entry, every element written
    std::vector<int> oc_off;       // per particle: offset of its per-orientation state (significant counts)
    long long total_oc = 0;
    std::vector<int> nd, np;  // per particle: directions, psi (an explicit list counts as nd = its length)
    HostList<double> or_prior, tr_prior;
    std::vector<long long> d2_off, n_entries;
    std::vector<float> xi2half, scale, scale_w;
    std::vector<double> old_off, prior_off;
    HostList<rlnk::CoarseTile> tiles;
    HostList<rlnk::WSeg> segs;  // coarse-weight segments (kSegEntries entries each)
    std::vector<int> seg_first;    // [B+1] first segment of each particle
    long long total_entries = 0;
    // per-particle scalars as uploaded, with the chunk capacity cap_B as stride so that their device addresses do
    // not depend on the chunk's size: xi2half, scale, scale_w; old offsets (x, y), prior offsets (x, y)
    std::vector<float> meta_f;
    std::vector<double> meta_d;
    // the uploaded arrays packed in page-locked memory (pack_chunk_host), offset and size of each; null: uploaded
    // from the vectors
    const unsigned char *pk = nullptr;
    size_t pk_off[16] = {}, pk_bytes[16] = {};
};

// Arrays of a chunk's upload, in packing order.
enum PackedArray {
    kPkOrOff, kPkOcOff, kPkOrCnt, kPkOrProj, kPkOrPrior, kPkOrFrame, kPkOrClass, kPkOrDir, kPkTrPrior, kPkTiles,
    kPkD2Off, kPkNent, kPkMetaF, kPkMetaD, kPkSegs, kPkSegFirst, kPkCount
};

// set_off/set_og (optional): explicit coarse orientation lists of the chunk's particles (global grid indices
// idir * n_psi + ipsi, offsets relative to set_off[0]); an empty list means the exhaustive grid.
// Two passes: sizes per particle, then every particle fills its own slices (OpenMP over particles).
static void build_chunk_host(Engine &e, int B, const rln_particle_meta *meta, ChunkHost &h,
                             const int *set_off = nullptr, const int *set_og = nullptr) {
    const rln_estep_config &cfg = e.cfg;
    const int K = e.K, NT = e.NT;
    // the lists of an earlier chunk are overwritten in place (no new allocation); every element of every list
    // below is written
    h.pk = nullptr;
    h.total_oc = h.total_entries = 0;
    h.B = B;
    h.or_off.resize(B);
    h.oc_off.resize(B);
    h.or_cnt.resize(B);
    h.nd.resize(B);
    h.np.resize(B);
    h.tr_prior.resize((size_t)B * NT);
    h.d2_off.resize(B);
    h.n_entries.resize(B);
    h.xi2half.resize(B);
    h.scale.resize(B);
    h.scale_w.resize(B);
    h.old_off.resize(2 * B);
    h.prior_off.resize(2 * B);
    const double px = cfg.pixel_size;
    const double s2 = cfg.sigma2_offset;
    const double offset_norm = 2. * kPi * sqrt(s2) * sqrt(s2);
    std::vector<LocalSet> sets;
    if (cfg.local_search) local_priors_batch(e.grid, cfg, meta, B, sets);
    const int n_dir = e.grid.n_dir, n_psi = e.grid.n_psi;
    auto n_set = [&](int b) { return set_off ? set_off[b + 1] - set_off[b] : 0; };
    // particles searched over the whole grid share one block of orientation entries at offset 0
    auto full_grid = [&](int b) { return !cfg.local_search && n_set(b) == 0; };
    bool any_full = false;
    for (int b = 0; b < B && !any_full; ++b) any_full = full_grid(b);
    const int n_grid = K * n_dir * n_psi;
    // pass 1: sizes
    for (int b = 0; b < B; ++b) {
        int nd, np;
        if (n_set(b) > 0) {
            nd = n_set(b);
            np = 1;
        } else if (cfg.local_search) {
            nd = (int)sets[b].dir_idx.size();
            np = (int)sets[b].psi_idx.size();
        } else {
            nd = n_dir;
            np = n_psi;
        }
        h.nd[b] = nd;
        h.np[b] = np;
        h.or_cnt[b] = K * nd * np;
    }
    long long tot_or = any_full ? n_grid : 0, d2 = 0, tot_oc = 0;
    std::vector<int> tile_off(B + 1, 0), seg_off(B + 1, 0);
    for (int b = 0; b < B; ++b) {
        h.oc_off[b] = (int)tot_oc;
        tot_oc += h.or_cnt[b];
        if (full_grid(b)) {
            h.or_off[b] = 0;
        } else {
            h.or_off[b] = (int)tot_or;
            tot_or += h.or_cnt[b];
        }
        h.n_entries[b] = (long long)h.or_cnt[b] * NT;
        h.d2_off[b] = d2;
        d2 += h.n_entries[b];
        tile_off[b + 1] = tile_off[b] + (h.or_cnt[b] + rlnk::kCoTO - 1) / rlnk::kCoTO;
        const long long ns = std::max(1LL, (h.n_entries[b] + rlnk::kSegEntries - 1) / rlnk::kSegEntries);
        seg_off[b + 1] = seg_off[b] + (int)ns;
    }
    h.total_entries = d2;
    h.total_oc = tot_oc;
    // Tile placement. Global searches enumerate the same grid for every particle: the coarse tiles are ordered by
    // orientation block first, so that blocks running together read the same projection rows (L2 hits) instead
    // of each particle streaming the whole grid from HBM. Tiles of particles searched over the whole grid (all of
    // them, or the exhaustive fallbacks among explicit lists) come first, tile j of the r-th such particle at
    // j * n_full + r; tiles of explicit lists after them in particle order. Local searches keep particle order.
    std::vector<int> tile_base(B), full_rank(B, -1);
    int n_full = 0;
    if (!cfg.local_search) {
        for (int b = 0; b < B; ++b)
            if (full_grid(b)) full_rank[b] = n_full++;
        const int t_full = (n_grid + rlnk::kCoTO - 1) / rlnk::kCoTO;
        int t = n_full * t_full;
        for (int b = 0; b < B; ++b)
            if (full_rank[b] < 0) {
                tile_base[b] = t;
                t += tile_off[b + 1] - tile_off[b];
            }
    } else {
        for (int b = 0; b < B; ++b) tile_base[b] = tile_off[b];
    }
    h.or_proj.resize(tot_or);
    h.or_prior.resize(tot_or);
    h.or_frame.resize(tot_or);
    h.or_class.resize(tot_or);
    h.or_dir.resize(tot_or);
    h.tiles.resize(tile_off[B]);
    h.segs.resize(seg_off[B]);
    h.seg_first.assign(seg_off.begin(), seg_off.end());
    if (any_full) {
        // the whole grid: orientation priors normalised by their mean over classes, directions and psi
        // (src/ml_optimiser.cpp:7790-7803), the same for every particle (pdf_dir_mean, summed in grid order)
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n_grid; ++q) {
            const int k = q / (n_dir * n_psi), og = q - k * (n_dir * n_psi), d = og / n_psi;
            const double po = e.pdf_direction[(size_t)k * n_dir + d];
            const bool on = po > 0. && e.pdf_class[k] > 0.;
            h.or_proj[q] = on ? k * e.n_or + og : -1;
            h.or_prior[q] = po / e.pdf_dir_mean;
            h.or_frame[q] = og * 8;
            h.or_class[q] = k;
            h.or_dir[q] = d;
        }
    }
    // pass 2: every particle fills its slices
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < B; ++b) {
        const rln_particle_meta &m = meta[b];
        const int base = h.or_off[b];
        int q = base;
        if (full_grid(b)) {
            // entries in the shared block
        } else if (n_set(b) > 0) {
            // explicit list: global-search priors, normalised by their mean over the whole grid so that weights
            // and sums stay those of the exhaustive search restricted to the list
            const int *og_list = set_og + (set_off[b] - set_off[0]);
            for (int k = 0; k < K; ++k)
                for (int i = 0; i < n_set(b); ++i, ++q) {
                    const int og = og_list[i], d = og / n_psi;
                    const double po = e.pdf_direction[(size_t)k * n_dir + d] / e.pdf_dir_mean;
                    const bool on = po > 0. && e.pdf_class[k] > 0.;
                    h.or_proj[q] = on ? k * e.n_or + og : -1;
                    h.or_prior[q] = po;
                    h.or_frame[q] = og * 8;
                    h.or_class[q] = k;
                    h.or_dir[q] = d;
                }
        } else {
            // orientation priors, normalised by their mean over classes, directions and psi
            // (src/ml_optimiser.cpp:7790-7803)
            const int ND = h.nd[b], NP = h.np[b];
            double mean = 0.;
            for (int k = 0; k < K; ++k)
                for (int id = 0; id < ND; ++id)
                    for (int ip = 0; ip < NP; ++ip, ++q) {
                        const int di = cfg.local_search ? sets[b].dir_idx[id] : id;
                        const int pi = cfg.local_search ? sets[b].psi_idx[ip] : ip;
                        const double po = cfg.local_search ? sets[b].dir_prior[id] * sets[b].psi_prior[ip]
                                                           : e.pdf_direction[(size_t)k * n_dir + di];
                        mean += po;
                        const int og = di * n_psi + pi;
                        const bool on = po > 0. && e.pdf_class[k] > 0.;
                        h.or_proj[q] = on ? k * e.n_or + og : -1;
                        h.or_prior[q] = po;
                        h.or_frame[q] = og * 8;
                        h.or_class[q] = k;
                        h.or_dir[q] = di;
                    }
            mean /= (double)((long long)K * ND * NP);
            if (mean != 0.)
                for (int i = 0; i < K * ND * NP; ++i) h.or_prior[base + i] /= mean;
        }
        // offset priors (src/ml_optimiser.cpp:7805-7872): offsets in pixels plus translations in Angstrom. The
        // Gaussian exp(td) factors into exp(x part) * exp(y part), evaluated once per distinct x and y value of the
        // translation lattice (within a few ulp of exp(td), against one exp per translation)
        double tmean = 0.;
        if (s2 < 0.0001) {
            for (int t = 0; t < NT; ++t) {
                const double ox = m.old_offset_x + e.grid.tx[t], oy = m.old_offset_y + e.grid.ty[t];
                double td = (ox - m.prior_offset_x) * (ox - m.prior_offset_x) / (-2. * s2);
                td += (oy - m.prior_offset_y) * (oy - m.prior_offset_y) / (-2. * s2);
                td *= px * px;
                const double pf = (td > 0.) ? 0. : 1.;
                h.tr_prior[(size_t)b * NT + t] = pf;
                tmean += pf;
            }
        } else {
            const size_t nux = e.tr_ux.size(), nuy = e.tr_uy.size();
            double fx_s[64], fy_s[64];
            std::vector<double> fx_v, fy_v;
            double *fx = fx_s, *fy = fy_s;
            if (nux > 64) fx = (fx_v.resize(nux), fx_v.data());
            if (nuy > 64) fy = (fy_v.resize(nuy), fy_v.data());
            for (size_t i = 0; i < nux; ++i) {
                const double d = m.old_offset_x + e.tr_ux[i] - m.prior_offset_x;
                fx[i] = exp(d * d / (-2. * s2) * px * px);
            }
            for (size_t i = 0; i < nuy; ++i) {
                const double d = m.old_offset_y + e.tr_uy[i] - m.prior_offset_y;
                fy[i] = exp(d * d / (-2. * s2) * px * px);
            }
            for (int t = 0; t < NT; ++t) {
                const double pf = fx[e.tr_cix[t]] * fy[e.tr_ciy[t]] / offset_norm;
                h.tr_prior[(size_t)b * NT + t] = pf;
                tmean += pf;
            }
        }
        tmean = tmean * K / (double)(K * NT);  // mean over classes x translations (same per class)
        if (tmean > 0.)
            for (int t = 0; t < NT; ++t) h.tr_prior[(size_t)b * NT + t] /= tmean;
        if (full_rank[b] >= 0) {
            for (int o = 0, j = 0; o < h.or_cnt[b]; o += rlnk::kCoTO, ++j)
                h.tiles[(size_t)j * n_full + full_rank[b]] = rlnk::CoarseTile{b, o};
        } else {
            int ti = tile_base[b];
            for (int o = 0; o < h.or_cnt[b]; o += rlnk::kCoTO) h.tiles[ti++] = rlnk::CoarseTile{b, o};
        }
        const long long n = h.n_entries[b];
        for (int sI = seg_off[b]; sI < seg_off[b + 1]; ++sI) {
            const long long qs = sI - seg_off[b];
            h.segs[sI] = rlnk::WSeg{b, 0, qs * rlnk::kSegEntries, std::min(n, (qs + 1) * rlnk::kSegEntries)};
        }
        h.xi2half[b] = (float)(m.highres_Xi2 / 2.);
        const double sc = cfg.do_scale_correction ? m.scale : 1.;
        h.scale[b] = (float)sc;
        h.scale_w[b] = (float)std::max(0.001, sc);
        h.old_off[2 * b] = m.old_offset_x;
        h.old_off[2 * b + 1] = m.old_offset_y;
        h.prior_off[2 * b] = m.prior_offset_x;
        h.prior_off[2 * b + 1] = m.prior_offset_y;
    }
    const int cB = e.cap_B;
    h.meta_f.assign(3 * (size_t)cB, 0.f);
    for (int b = 0; b < B; ++b) {
        h.meta_f[b] = h.xi2half[b];
        h.meta_f[cB + b] = h.scale[b];
        h.meta_f[2 * cB + b] = h.scale_w[b];
    }
    h.meta_d.assign(5 * (size_t)cB, 0.);
    for (int b = 0; b < 2 * B; ++b) {
        h.meta_d[b] = h.old_off[b];
        h.meta_d[2 * cB + b] = h.prior_off[b];
    }
    for (int b = 0; b < B; ++b) h.meta_d[4 * (size_t)cB + b] = meta[b].highres_Xi2 / 2.;
}

// Host array `which` (PackedArray) of a chunk's upload: address and bytes.
static std::pair<const void *, size_t> chunk_array(const ChunkHost &h, int which) {
    auto src = [](const auto &v) { return std::make_pair((const void *)v.data(), v.size() * sizeof(v[0])); };
    switch (which) {
        case kPkOrOff: return src(h.or_off);
        case kPkOcOff: return src(h.oc_off);
        case kPkOrCnt: return src(h.or_cnt);
        case kPkOrProj: return src(h.or_proj);
        case kPkOrPrior: return src(h.or_prior);
        case kPkOrFrame: return src(h.or_frame);
        case kPkOrClass: return src(h.or_class);
        case kPkOrDir: return src(h.or_dir);
        case kPkTrPrior: return src(h.tr_prior);
        case kPkTiles: return src(h.tiles);
        case kPkD2Off: return src(h.d2_off);
        case kPkNent: return src(h.n_entries);
        case kPkMetaF: return src(h.meta_f);
        case kPkMetaD: return src(h.meta_d);
        case kPkSegs: return src(h.segs);
        default: return src(h.seg_first);
    }
}

// Packs a chunk's upload into page-locked memory, each array 256-byte aligned, so that run_chunk uploads it with
// copies asynchronous to the host. For chunks built on the second host thread while the device runs the previous
// chunk; the first chunk of a run is uploaded from its vectors, as packing it would delay the device's start.
static int pack_chunk_host(Engine &e, ChunkHost &h, PinnedBuf &pb) {
    size_t tot = 0;
    for (int i = 0; i < kPkCount; ++i) {
        h.pk_off[i] = tot;
        h.pk_bytes[i] = chunk_array(h, i).second;
        tot += (h.pk_bytes[i] + 255) & ~(size_t)255;
    }
    HIP_TRY(hipSetDevice(e.device));
    RLN_TRY(pb.ensure(tot));
    unsigned char *dst = static_cast<unsigned char *>(pb.ptr);
    for (int i = 0; i < kPkCount; ++i) {
        const std::pair<const void *, size_t> a = chunk_array(h, i);
        if (a.second) std::memcpy(dst + h.pk_off[i], a.first, a.second);
    }
    h.pk = dst;
    return RLN_OK;
}

// One array of a chunk's upload to its device buffer (grown to at least 256 bytes), from the page-locked
// package when there is one, else from the host vector.
static int upload_packed(DevBuf &buf, const ChunkHost &h, int which, hipStream_t st) {
    const std::pair<const void *, size_t> a = chunk_array(h, which);
    RLN_TRY(buf.ensure(std::max<size_t>(256, a.second)));
    if (a.second) {
        const void *src = h.pk ? (const void *)(h.pk + h.pk_off[which]) : a.first;
        HIP_TRY(hipMemcpyAsync(buf.ptr, src, a.second, hipMemcpyHostToDevice, st));
    }
    return RLN_OK;
}

// ------------------------------------------------------------------------------------------------
// Fine pass planned on the device (persistent kernels, one captured graph)
// ------------------------------------------------------------------------------------------------
static int ensure_fine_capacity(Engine &e) {
    const int cB = e.cap_B;
    if (e.need_pairs == 0) e.need_pairs = 64LL * cB;
    if (e.need_jobs == 0) e.need_jobs = 16 * cB;
    if (e.need_small == 0) e.need_small = 16 * cB;
    if (e.need_wide == 0) e.need_wide = 16 * cB;
    RLN_TRY(e.d_pairs.ensure(sizeof(int) * e.need_pairs));
    RLN_TRY(e.d_jobs.ensure(sizeof(rlnk::FineJob) * e.need_jobs));
    RLN_TRY(e.d_job_order.ensure(sizeof(int) * e.need_jobs));
    RLN_TRY(e.d_items_s.ensure(sizeof(int2) * e.need_small));
    RLN_TRY(e.d_items_w.ensure(sizeof(int2) * e.need_wide));
    RLN_TRY(e.d_fdiff2.ensure(sizeof(double) * 32 * e.need_pairs));
    RLN_TRY(e.d_fweight.ensure(sizeof(double) * 32 * e.need_pairs));
    RLN_TRY(e.d_npairs.ensure(sizeof(int) * 2 * cB));
    RLN_TRY(e.d_pair_off.ensure(sizeof(long long) * cB));
    RLN_TRY(e.d_job_off.ensure(sizeof(int) * cB));
    RLN_TRY(e.d_fine_n.ensure(sizeof(long long) * cB));
    RLN_TRY(e.d_foff.ensure(sizeof(long long) * cB));
    RLN_TRY(e.d_stats_f.ensure(sizeof(rlnk::PassStats) * cB));
    RLN_TRY(e.d_shell_acc.ensure(sizeof(double) * e.nshell * cB));
    RLN_TRY(e.d_shell_tot.ensure(sizeof(double) * e.nshell));
    RLN_TRY(e.d_part_acc.ensure(sizeof(double) * 8 * cB));
    RLN_TRY(e.d_pdfc_acc.ensure(sizeof(double) * e.K * cB));
    RLN_TRY(e.d_plan.ensure(sizeof(rlnk::FinePlan)));
    RLN_TRY(e.d_cp.ensure(sizeof(rlnk::ChunkPtrs)));
    return RLN_OK;
}

template <int kTuning>
static void launch_fine_shape(int shape, dim3 grid, hipStream_t st, const rlnk::FineArgs &fa) {
    switch (shape) {
        case 1: hipLaunchKernelGGL((rlnk::k_fine_diff2<kTuning, 1, 2>), grid, dim3(rlnk::kBlock), 0, st, fa); break;
        case 2: hipLaunchKernelGGL((rlnk::k_fine_diff2<kTuning, 2, 1>), grid, dim3(rlnk::kBlock), 0, st, fa); break;
        case 3: hipLaunchKernelGGL((rlnk::k_fine_diff2<kTuning, 1, 1>), grid, dim3(rlnk::kBlock), 0, st, fa); break;
        default: hipLaunchKernelGGL((rlnk::k_fine_diff2<kTuning, 2, 2>), grid, dim3(rlnk::kBlock), 0, st, fa); break;
    }
}

// All kernel arguments of the fine pass of a chunk; the graph is reused while they are unchanged.
struct FineLaunch {
    const int *np, *nj;
    const rlnk::ChunkPtrs *cp;
    long long cap_pairs;
    int cap_jobs, cap_small, cap_wide, grid_B;
    long long *pair_off, *fine_n, *foff;
    int *job_off, *pairs;
    rlnk::FinePlan *plan;
    const unsigned char *mask;
    const long long *d2off;
    const int *or_cnt, *oc_off, *ocnt;
    int NT;
    rlnk::FineJob *jobs;
    int *job_order;
    const int *or_off, *or_frame, *or_class;
    int n_or_grid, key_shift, sorted;  // job keys (class, coarse orientation) >> key_shift; sorted: orientation order
    int2 *items_s, *items_w;
    const double *fbase;
    int shape;  // work-item shape (index into rlnk::kFineShapes)
    rlnk::FineArgs fa;
    rlnk::WeightArgs wf;
    rlnk::WsumArgs ws;
    double *shell, *part, *pdfc;
    long long n_shell, n_part, n_pdfc;
    int grid_fine, grid_wsum;
};

// part 0: plan, job lists and work items; part 1: fine diff2; part 2: fine weights; part 3: weighted sums and
// backprojection. Each part is captured into its own graph so that stage events sit between graph launches.
static void launch_plan(Engine &e, const FineLaunch &L) {
    hipStream_t st = e.stream;
    hipLaunchKernelGGL(rlnk::k_fine_plan, dim3(1), dim3(rlnk::kWBlock), 0, st, L.np, L.nj, L.cp, L.cap_pairs,
                       L.cap_jobs, L.pair_off, L.job_off, L.fine_n, L.foff, L.plan);
    hipLaunchKernelGGL(rlnk::k_write_jobs, dim3(L.grid_B), dim3(rlnk::kWBlock), 0, st, L.mask, L.d2off, L.or_cnt,
                       L.oc_off, L.ocnt, L.NT,
                       (const long long *)L.pair_off, (const int *)L.job_off, L.pairs, L.jobs,
                       (const rlnk::FinePlan *)L.plan);
    const rlnk::FineShape sh = rlnk::kFineShapes[L.shape];
    hipLaunchKernelGGL(rlnk::k_fine_items, dim3(1), dim3(rlnk::kWBlock), 0, st, (const rlnk::FineJob *)L.jobs,
                       L.or_off, L.or_frame, L.or_class, L.n_or_grid, L.key_shift, L.sorted, sh.tc, 4 * sh.cw,
                       L.cap_small, L.cap_wide, L.job_order, L.items_s, L.items_w, L.plan);
}

static int launch_fine(Engine &e, const FineLaunch &L, int part) {
    hipStream_t st = e.stream;
    switch (part) {
        case 0:
            launch_plan(e, L);
            return launch_check("fine plan");
        case 1:
            hipLaunchKernelGGL(rlnk::k_fine_init, dim3(L.grid_B), dim3(rlnk::kBlock), 0, st, L.np, L.pair_off, L.fbase,
                               (const rlnk::FinePlan *)L.plan, L.fa.diff2);
            launch_fine_shape<0>(L.shape, dim3(L.grid_fine), st, L.fa);
            return launch_check("fine pass");
        case 2:
            hipLaunchKernelGGL(rlnk::k_weights, dim3(L.grid_B), dim3(rlnk::kWBlock), 0, st, L.wf);
            return launch_check("fine weights");
        default:
            hipLaunchKernelGGL(rlnk::k_zero_d, dim3(64), dim3(rlnk::kBlock), 0, st, L.shell, L.n_shell);
            hipLaunchKernelGGL(rlnk::k_zero_d, dim3(16), dim3(rlnk::kBlock), 0, st, L.part, L.n_part);
            hipLaunchKernelGGL(rlnk::k_zero_d, dim3(16), dim3(rlnk::kBlock), 0, st, L.pdfc, L.n_pdfc);
            hipLaunchKernelGGL(rlnk::k_wsum_bp, dim3(L.grid_wsum), dim3(rlnk::kBlock), 0, st, L.ws);
            return launch_check("weighted sums");
    }
}

// Work-item shape, unit order, work-unit size and persistent grid of the fine diff2 kernel, timed on the first chunk
// with enough work items, one parameter after the other at the best values found so far: the work-item shape
// (coarse translations per small item and per wave of a wide item, which set the register tiles and the projections
// shared per item), the order of the units (jobs in particle order or in orientation order, units band-major or
// item-major; kFineOrders), the band count (pixels per unit: all sampled pixels, or the radial shells of
// 1/2 .. 1/32 of them) and the grid (512 .. 4096 blocks or the shape's resident-block count); then the other shapes
// at every grid, since the best grid depends on the shape. The plan's items are rebuilt for each shape and order. The fastest combination is kept for the engine's life. The kernel only adds into
// fine diff2, which the production pass initialises again, so repeating it is harmless; its unit counter is reset
// before every candidate. tune_ms keeps every candidate's time ([shape][order][band][grid], -1 untried) for
// rln_estep_fine_tuning.
// A candidate replaces the best one found so far only when it is at least 1% faster, so that candidates within the
// timing noise of each other do not alternate between engines.
constexpr float kTuneMargin = 0.99f;

static int tune_fine(Engine &e, FineLaunch L) {
    hipStream_t st = e.stream;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    const int n_grid = 5;
    auto grid_of = [&](int sh, int gi) {
        static const int g[4] = {512, 1024, 2048, 4096};
        const int res = e.resident_shape[sh] > 0 ? e.resident_shape[sh] * e.n_cu : 1024;
        return gi < 4 ? g[gi] : res;
    };
    e.tune_ms.assign((size_t)rlnk::kNumFineShapes * kNumFineOrders * kNumBandChoices * n_grid, -1.);
    // plan of a shape and order; ok false when it does not fit the capacities (the candidate is skipped) or is too
    // small to tune
    auto plan_for = [&](int sh, int oi, bool &ok) -> int {
        FineLaunch P = L;
        P.shape = sh;
        P.sorted = kFineOrders[oi].sorted;
        launch_plan(e, P);
        RLN_TRY(launch_check("fine plan"));
        rlnk::FinePlan plan{};
        HIP_TRY(hipMemcpyAsync(&plan, L.plan, sizeof(plan), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        ok = !plan.overflow && plan.n_small + plan.n_wide >= 512;
        return RLN_OK;
    };
    auto time_one = [&](int sh, int oi, int ci, int gi, float &ms) -> int {
        rlnk::FineArgs fa = L.fa;
        fa.n_bands = kBandChoices[ci];
        fa.item_major = kFineOrders[oi].item_major;
        fa.xcd_queues = kFineOrders[oi].xcd_queues;
        fa.pix = e.d_band_pix.as<const int2>() + (size_t)ci * e.pf_sampled;
        hipLaunchKernelGGL(rlnk::k_reset_items, dim3(1), dim3(64), 0, st, L.plan);
        HIP_TRY(hipEventRecord(a, st));
        launch_fine_shape<1>(sh, dim3(grid_of(sh, gi)), st, fa);
        HIP_TRY(hipEventRecord(b, st));
        HIP_TRY(hipEventSynchronize(b));
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        e.tune_ms[(((size_t)sh * kNumFineOrders + oi) * kNumBandChoices + ci) * n_grid + gi] = ms;
        return RLN_OK;
    };
    int best_o = e.order_fixed >= 0 ? e.order_fixed : 0;
    bool ok = false;
    RLN_TRY(plan_for(0, best_o, ok));
    if (!ok) {  // retried with more room, or too small to be worth tuning
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return RLN_OK;
    }
    // the first band count with at least 8 bands of >= 256 pixels, or the largest below
    int best_c = 0;
    for (int ci = 0; ci < kNumBandChoices; ++ci)
        if (e.pf_sampled / kBandChoices[ci] >= 256 && kBandChoices[ci] <= 8) best_c = ci;
    int best_s = e.shape_fixed >= 0 ? e.shape_fixed : 0, best_g = n_grid - 1;
    float best = 1e30f;
    for (int sh = 0; sh < rlnk::kNumFineShapes; ++sh) {
        if (e.shape_fixed >= 0 && sh != e.shape_fixed) continue;
        if (sh > 0 || e.shape_fixed >= 0) {
            RLN_TRY(plan_for(sh, best_o, ok));
            if (!ok) continue;
        }
        float ms = 0.f;
        RLN_TRY(time_one(sh, best_o, best_c, best_g, ms));
        if (ms < kTuneMargin * best) {
            best = ms;
            best_s = sh;
        }
    }
    if (e.order_fixed < 0) {
        const int o0 = best_o;
        for (int oi = 0; oi < kNumFineOrders; ++oi) {
            if (oi == o0) continue;
            RLN_TRY(plan_for(best_s, oi, ok));
            if (!ok) continue;
            float ms = 0.f;
            RLN_TRY(time_one(best_s, oi, best_c, best_g, ms));
            if (ms < kTuneMargin * best) {
                best = ms;
                best_o = oi;
            }
        }
    }
    RLN_TRY(plan_for(best_s, best_o, ok));
    for (int ci = 0; ci < kNumBandChoices; ++ci) {
        if (ci == best_c || (kBandChoices[ci] > 1 && e.pf_sampled / kBandChoices[ci] < 256)) continue;
        float ms = 0.f;
        RLN_TRY(time_one(best_s, best_o, ci, best_g, ms));
        if (ms < kTuneMargin * best) {
            best = ms;
            best_c = ci;
        }
    }
    for (int gi = 0; gi + 1 < n_grid; ++gi) {
        float ms = 0.f;
        RLN_TRY(time_one(best_s, best_o, best_c, gi, ms));
        if (ms < kTuneMargin * best) {
            best = ms;
            best_g = gi;
        }
    }
    // the other shapes again, at the order and band count kept and every grid: a shape's best grid can differ
    if (e.shape_fixed < 0) {
        const int s0 = best_s;
        for (int sh = 0; sh < rlnk::kNumFineShapes; ++sh) {
            if (sh == s0) continue;
            RLN_TRY(plan_for(sh, best_o, ok));
            if (!ok) continue;
            for (int gi = 0; gi < n_grid; ++gi) {
                float ms = 0.f;
                RLN_TRY(time_one(sh, best_o, best_c, gi, ms));
                if (ms < kTuneMargin * best) {
                    best = ms;
                    best_s = sh;
                    best_g = gi;
                }
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    RLN_TRY(launch_check("fine tuning"));
    e.fine_shape = best_s;
    e.fine_order = best_o;
    e.band_choice = best_c;
    e.grid_fine = grid_of(best_s, best_g);
    e.tuned = true;
    return RLN_OK;
}

static void drop_fine_graphs(Engine &e) {
    for (int q = 0; q < kFineGraphs; ++q) {
        if (e.fine_exec[q]) (void)hipGraphExecDestroy(e.fine_exec[q]);
        if (e.fine_graph[q]) (void)hipGraphDestroy(e.fine_graph[q]);
        e.fine_exec[q] = nullptr;
        e.fine_graph[q] = nullptr;
    }
}

static int run_fine_stage(Engine &e, const FineLaunch &L0) {
    FineLaunch L = L0;
    if (!e.tuned) {
        RLN_TRY(tune_fine(e, L));
        L.grid_fine = e.grid_fine;
        L.shape = e.fine_shape;
        L.sorted = kFineOrders[e.fine_order].sorted;
        L.fa.item_major = kFineOrders[e.fine_order].item_major;
        L.fa.xcd_queues = kFineOrders[e.fine_order].xcd_queues;
        L.fa.n_bands = kBandChoices[e.band_choice];
        L.ws.n_bands = kBandChoices[wsum_band_choice(e)];
        L.ws.wpix = e.d_wsum_pix.as<const int2>() + (size_t)wsum_band_choice(e) * e.n_wsum_pix;
        L.fa.pix = e.d_band_pix.as<const int2>() + (size_t)e.band_choice * e.pf_sampled;
    }
    if (!e.graph_off) {
        const unsigned char *kb = reinterpret_cast<const unsigned char *>(&L);
        std::vector<unsigned char> key(kb, kb + sizeof(L));
        if (!e.fine_exec[0] || key != e.fine_key) {
            drop_fine_graphs(e);
            bool ok = true;
            for (int q = 0; q < kFineGraphs && ok; ++q) {
                ok = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
                if (!ok) break;
                const int rc = launch_fine(e, L, q);
                hipGraph_t g = nullptr;
                ok = hipStreamEndCapture(e.stream, &g) == hipSuccess && rc == RLN_OK && g;
                if (ok) ok = hipGraphInstantiate(&e.fine_exec[q], g, nullptr, nullptr, 0) == hipSuccess;
                if (ok) e.fine_graph[q] = g;
                else if (g) (void)hipGraphDestroy(g);
            }
            (void)hipGetLastError();
            if (!ok) {
                drop_fine_graphs(e);
                e.graph_off = true;  // capture unsupported here: launch directly from now on
            } else {
                e.fine_key = key;
            }
        }
    }
    static const Stage before[kFineGraphs] = {kStJobs, kStFine, kStFineW, kStWsum};
    for (int q = 0; q < kFineGraphs; ++q) {
        HIP_TRY(hipEventRecord(e.ev[before[q]], e.stream));
        if (e.fine_exec[q]) {
            HIP_TRY(hipGraphLaunch(e.fine_exec[q], e.stream));
            e.counters[12] += 1;
        } else {
            RLN_TRY(launch_fine(e, L, q));
        }
    }
    return RLN_OK;
}

template <int NTT>
static void launch_mfma_t(Engine &e, unsigned n_tiles, const float *xi2, const long long *d2off) {
    hipLaunchKernelGGL(rlnk::k_coarse_mfma<NTT>, dim3((n_tiles + rlnk::kMfG - 1) / rlnk::kMfG),
                       dim3(rlnk::kMfBlock), 0, e.stream, e.d_tiles.as<const rlnk::CoarseTile>(), (int)n_tiles,
                       e.d_or_off.as<const int>(), e.d_or_cnt.as<const int>(),
                       e.d_or_proj.as<const int>(), e.d_pf.as<const float2>(), e.d_wc.as<const float2>(), e.d_ctf2.as<const float>(), e.d_ynorm.as<const double>(),
                       e.d_ehi.as<const unsigned int>(), e.d_elo.as<const unsigned int>(), xi2, d2off, e.NT, e.Kp,
                       e.d_zero_row.as<const float2>(), e.d_diff2.as<float>(), e.d_wpart.as<rlnk::WPart>());
}

static int launch_coarse_mfma(Engine &e, unsigned n_tiles, const float *xi2, const long long *d2off) {
    const int ntt = (e.NT + 15) / 16;
    switch (ntt) {
#define RLN_MFMA_CASE(N)                          \
    case N:                                       \
        launch_mfma_t<N>(e, n_tiles, xi2, d2off); \
        break;
        RLN_MFMA_CASE(1) RLN_MFMA_CASE(2) RLN_MFMA_CASE(3) RLN_MFMA_CASE(4) RLN_MFMA_CASE(5) RLN_MFMA_CASE(6)
        RLN_MFMA_CASE(7) RLN_MFMA_CASE(8) RLN_MFMA_CASE(9) RLN_MFMA_CASE(10) RLN_MFMA_CASE(11) RLN_MFMA_CASE(12)
        RLN_MFMA_CASE(13) RLN_MFMA_CASE(14) RLN_MFMA_CASE(15) RLN_MFMA_CASE(16)
#undef RLN_MFMA_CASE
        default:
            RLN_FAIL(RLN_ERR_ARG, "split-BF16 coarse pass supports at most 256 translations (%d)", e.NT);
    }
    return launch_check("k_coarse_mfma");
}

// The host lists of a chunk (h) are built beforehand, for the chunk after the first on a second host thread
// while the device runs the previous chunk.
static int run_chunk(Engine &e, int B, const float *fimg, const float *fimgn, const float *fctf, int on_device,
                     const rln_particle_meta *meta, const ChunkHost &h, rln_particle_result *results, rln_wsum *wsum) {
    const rln_estep_config &cfg = e.cfg;
    const int K = e.K, NT = e.NT;
    hipStream_t st = e.stream;
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t0) {
        return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    };
    for (int b = 0; b < B; ++b)
        if (meta[b].scale > 10000. && cfg.do_scale_correction)
            RLN_FAIL(RLN_ERR_ARG, "rlnMicrographScaleCorrection is very high (particle %d)", b);
    const size_t Pf = e.igf.P, Pc = e.igc.P;
    HIP_TRY(hipEventRecord(e.ev[kStPrep], st));
    // inputs
    const float2 *dX, *dXn;
    const float *dC;
    if (on_device) {
        dX = reinterpret_cast<const float2 *>(fimg);
        dXn = reinterpret_cast<const float2 *>(fimgn);
        dC = fctf;
    } else {
        RLN_TRY(e.d_fimg.ensure(sizeof(float2) * Pf * B));
        RLN_TRY(e.d_fctf.ensure(sizeof(float) * Pf * B));
        HIP_TRY(hipMemcpyAsync(e.d_fimg.ptr, fimg, sizeof(float2) * Pf * B, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(e.d_fctf.ptr, fctf, sizeof(float) * Pf * B, hipMemcpyHostToDevice, st));
        if (fimgn != fimg) {
            RLN_TRY(e.d_fimgn.ensure(sizeof(float2) * Pf * B));
            HIP_TRY(hipMemcpyAsync(e.d_fimgn.ptr, fimgn, sizeof(float2) * Pf * B, hipMemcpyHostToDevice, st));
            dXn = e.d_fimgn.as<const float2>();
        } else {
            dXn = e.d_fimg.as<const float2>();
        }
        dX = e.d_fimg.as<const float2>();
        dC = e.d_fctf.as<const float>();
    }
    // the chunk's lists, packed in page-locked memory when they were built: asynchronous copies
    RLN_TRY(upload_packed(e.d_or_off, h, kPkOrOff, st));
    RLN_TRY(upload_packed(e.d_oc_off, h, kPkOcOff, st));
    RLN_TRY(upload_packed(e.d_or_cnt, h, kPkOrCnt, st));
    RLN_TRY(upload_packed(e.d_or_proj, h, kPkOrProj, st));
    RLN_TRY(upload_packed(e.d_or_prior, h, kPkOrPrior, st));
    RLN_TRY(upload_packed(e.d_or_frame, h, kPkOrFrame, st));
    RLN_TRY(upload_packed(e.d_or_class, h, kPkOrClass, st));
    RLN_TRY(upload_packed(e.d_or_dir, h, kPkOrDir, st));
    RLN_TRY(upload_packed(e.d_tr_prior, h, kPkTrPrior, st));
    RLN_TRY(upload_packed(e.d_tiles, h, kPkTiles, st));
    RLN_TRY(upload_packed(e.d_meta_ll, h, kPkD2Off, st));
    DevBuf &d_d2off = e.d_meta_ll;
    DevBuf &d_nent_tmp = e.d_nent;  // entries per particle
    RLN_TRY(upload_packed(d_nent_tmp, h, kPkNent, st));
    // per-particle scalars with the chunk capacity as stride
    const int cB = e.cap_B;
    RLN_TRY(upload_packed(e.d_meta_f, h, kPkMetaF, st));
    RLN_TRY(upload_packed(e.d_meta_d, h, kPkMetaD, st));
    const float *d_xi2 = e.d_meta_f.as<const float>();
    const float *d_scale = d_xi2 + cB;
    const float *d_scale_w = d_xi2 + 2 * cB;
    const double *d_old = e.d_meta_d.as<const double>();
    const double *d_prior = d_old + 2 * cB;
    const double *d_xi2d = d_old + 4 * cB;  // highres_Xi2 / 2 in FP64

    // coarse: scaled images, shifted images
    const bool split = cfg.precision == RLN_PREC_SPLIT_BF16;
    if (split) {
        const size_t nc = (size_t)B * (e.Kp / 2);
        RLN_TRY(e.d_wc.ensure(nc * sizeof(float2)));
        RLN_TRY(e.d_ctf2.ensure(nc * sizeof(float)));
        RLN_TRY(e.d_ynorm.ensure(sizeof(double) * B));
        HIP_TRY(hipMemsetAsync(e.d_ynorm.ptr, 0, sizeof(double) * B, st));
        const int nthr = std::max(e.Kp / 2, (int)Pc);
        hipLaunchKernelGGL(rlnk::k_prep_coarse_split, dim3((nthr + rlnk::kBlock - 1) / rlnk::kBlock, B),
                           dim3(rlnk::kBlock), 0, st, dX, dC, d_scale, e.sf, (int)Pf, e.igc, e.d_q_c.as<const float>(),
                           e.d_grp.as<const int2>(), e.G, e.Kp, cfg.do_ctf_correction, e.d_wc.as<float2>(),
                           e.d_ctf2.as<float>(), e.d_ynorm.as<double>());
        RLN_TRY(launch_check("k_prep_coarse_split"));
    } else {
        RLN_TRY(e.d_ctfq.ensure(sizeof(float) * Pc * B));
        RLN_TRY(e.d_yq.ensure(sizeof(float2) * Pc * NT * B));
        hipLaunchKernelGGL(rlnk::k_prep_coarse, dim3((Pc + rlnk::kBlock - 1) / rlnk::kBlock, NT, B),
                           dim3(rlnk::kBlock), 0, st, dX, dC, d_scale, e.sf, (int)Pf, e.igc,
                           e.d_q_c.as<const float>(), e.d_ex_c.as<const float2>(), e.d_ey_c.as<const float2>(), NT,
                           cfg.do_ctf_correction, e.d_ctfq.as<float>(), e.d_yq.as<float2>());
        RLN_TRY(launch_check("k_prep_coarse"));
    }
    HIP_TRY(hipEventRecord(e.ev[kStCoarse], st));
    RLN_TRY(e.d_diff2.ensure(sizeof(float) * std::max<long long>(1, h.total_entries)));
    RLN_TRY(e.d_weight.ensure(sizeof(double) * std::max<long long>(1, h.total_entries)));
    RLN_TRY(e.d_mask.ensure(std::max<long long>(1, h.total_entries)));
    // per-particle selection state; the split coarse kernel folds the minimum diff2 into it as it writes
    const unsigned gB = (unsigned)((B + rlnk::kBlock - 1) / rlnk::kBlock);
    RLN_TRY(e.d_wpart.ensure(sizeof(rlnk::WPart) * B));
    hipLaunchKernelGGL(rlnk::k_wpart_init, dim3(gB), dim3(rlnk::kBlock), 0, st, e.d_wpart.as<rlnk::WPart>(), B);
    if (split && !h.tiles.empty()) {
        RLN_TRY(launch_coarse_mfma(e, (unsigned)h.tiles.size(), d_xi2, d_d2off.as<const long long>()));
    } else if (!h.tiles.empty()) {
        hipLaunchKernelGGL(rlnk::k_coarse_diff2, dim3((unsigned)h.tiles.size(), (NT + rlnk::kCoTT - 1) / rlnk::kCoTT),
                           dim3(rlnk::kBlock), 0, st, e.d_tiles.as<const rlnk::CoarseTile>(), e.d_or_off.as<const int>(),
                           e.d_or_cnt.as<const int>(), e.d_or_proj.as<const int>(), e.d_proj_c.as<const float2>(),
                           e.d_ctfq.as<const float>(), e.d_yq.as<const float2>(), d_xi2, d_d2off.as<const long long>(),
                           NT, (int)Pc, e.d_diff2.as<float>());
        RLN_TRY(launch_check("k_coarse_diff2"));
    }
    HIP_TRY(hipEventRecord(e.ev[kStCoarseW], st));
    RLN_TRY(e.d_stats_c.ensure(sizeof(rlnk::PassStats) * B));
    rlnk::WeightArgs wa{};
    wa.diff2 = e.d_diff2.as<const float>();
    wa.d2_off = d_d2off.as<const long long>();
    wa.n_entries = d_nent_tmp.as<const long long>();
    wa.or_prior = e.d_or_prior.as<const double>();
    wa.or_off = e.d_or_off.as<const int>();
    wa.oc_off = e.d_oc_off.as<const int>();
    wa.tr_prior = e.d_tr_prior.as<const double>();
    wa.NT = NT;
    wa.fine = 0;
    wa.adaptive_fraction = cfg.adaptive_fraction;
    wa.maxsig = cfg.maximum_significants;
    wa.weight = e.d_weight.as<double>();
    wa.mask = e.d_mask.as<unsigned char>();
    wa.stats = e.d_stats_c.as<rlnk::PassStats>();
    // coarse weights over segments of kSegEntries poses (multi-block per particle)
    {
        const unsigned nseg = (unsigned)h.segs.size();
        RLN_TRY(upload_packed(e.d_wsegs, h, kPkSegs, st));
        RLN_TRY(upload_packed(e.d_seg_first, h, kPkSegFirst, st));
        RLN_TRY(e.d_hcnt.ensure(sizeof(unsigned int) * 256 * B));
        RLN_TRY(e.d_hsum.ensure(sizeof(double) * 256 * B));
        RLN_TRY(e.d_wcand.ensure(sizeof(double) * rlnk::kGather * B));
        RLN_TRY(e.d_segbest.ensure(sizeof(rlnk::WSegBest) * nseg));
        HIP_TRY(hipMemsetAsync(e.d_hcnt.ptr, 0, sizeof(unsigned int) * 256 * B, st));
        // per-orientation significant counts and per-particle job counts of the fine pass
        RLN_TRY(ensure_fine_capacity(e));
        RLN_TRY(e.d_ocnt.ensure(sizeof(int) * std::max<long long>(1, h.total_oc)));
        HIP_TRY(hipMemsetAsync(e.d_ocnt.ptr, 0, sizeof(int) * std::max<long long>(1, h.total_oc), st));
        HIP_TRY(hipMemsetAsync(e.d_npairs.as<int>() + cB, 0, sizeof(int) * B, st));
        HIP_TRY(hipMemsetAsync(e.d_hsum.ptr, 0, sizeof(double) * 256 * B, st));
        const rlnk::WSeg *sg = e.d_wsegs.as<const rlnk::WSeg>();
        rlnk::WPart *wp = e.d_wpart.as<rlnk::WPart>();
        unsigned int *hc = e.d_hcnt.as<unsigned int>();
        double *hs = e.d_hsum.as<double>();
        if (!split) hipLaunchKernelGGL(rlnk::k_wseg_min, dim3(nseg), dim3(rlnk::kBlock), 0, st, wa, sg, wp);
        hipLaunchKernelGGL(rlnk::k_wseg_exp, dim3(nseg), dim3(rlnk::kBlock), 0, st, wa, sg, wp, hc, hs);
        hipLaunchKernelGGL(rlnk::k_wsel_start, dim3(gB), dim3(rlnk::kBlock), 0, st, wa, wp, B);
        for (int shift = 56; shift >= 0; shift -= 8) {
            if (shift < 56)  // the top byte's histogram came with the weights
                hipLaunchKernelGGL(rlnk::k_wseg_hist, dim3(nseg), dim3(rlnk::kBlock), 0, st, wa, sg,
                                   (const rlnk::WPart *)wp, shift, hc, hs);
            hipLaunchKernelGGL(rlnk::k_wsel_pick, dim3(gB), dim3(rlnk::kBlock), 0, st, wp, B, shift, hc, hs);
        }
        hipLaunchKernelGGL(rlnk::k_wseg_gather, dim3(nseg), dim3(rlnk::kBlock), 0, st, wa, sg, wp,
                           e.d_wcand.as<double>());
        hipLaunchKernelGGL(rlnk::k_wsel_final, dim3(B), dim3(rlnk::kWBlock), 0, st, wa, wp,
                           e.d_wcand.as<const double>());
        if (wa.maxsig > 0) hipLaunchKernelGGL(rlnk::k_wsel_maxsig, dim3(B), dim3(rlnk::kWBlock), 0, st, wa, wp);
        hipLaunchKernelGGL(rlnk::k_wseg_mask, dim3(nseg), dim3(rlnk::kBlock), 0, st, wa, sg, wp,
                           e.d_segbest.as<rlnk::WSegBest>(), e.d_ocnt.as<int>(), e.d_npairs.as<int>() + cB);
        hipLaunchKernelGGL(rlnk::k_wsel_stats, dim3(gB), dim3(rlnk::kBlock), 0, st, (const rlnk::WPart *)wp,
                           e.d_seg_first.as<const int>(), e.d_segbest.as<const rlnk::WSegBest>(), B, 0,
                           e.d_stats_c.as<rlnk::PassStats>(), e.d_npairs.as<int>());
        RLN_TRY(launch_check("coarse weights"));
    }
    HIP_TRY(hipEventRecord(e.ev[kStJobs], st));
    // ---------------------------------------------------------------- fine pass, planned on the device
    // per-particle counts; everything after runs from device-side counts (FinePlan) without host round trips,
    // as persistent kernels, captured once into a graph and replayed for every chunk
    // (per-particle pair and job counts were counted by k_wseg_mask / k_wsel_stats)
    e.cp_host.fimg = dX;
    e.cp_host.fimg_nomask = dXn;
    e.cp_host.fctf = dC;
    e.cp_host.B = B;
    HIP_TRY(hipMemcpyAsync(e.d_cp.ptr, &e.cp_host, sizeof(rlnk::ChunkPtrs), hipMemcpyHostToDevice, st));
    RLN_TRY(e.d_fbase.ensure(sizeof(double) * cB));
    hipLaunchKernelGGL(rlnk::k_fine_base, dim3(B), dim3(rlnk::kBlock), 0, st, e.d_cp.as<const rlnk::ChunkPtrs>(),
                       e.d_q_f.as<const float>(), (int)Pf, d_xi2d, e.d_fbase.as<double>());
    RLN_TRY(launch_check("k_fine_base"));
    rlnk::FinePlan plan{};
    for (int attempt = 0;; ++attempt) {
        RLN_TRY(ensure_fine_capacity(e));
        FineLaunch L;
        std::memset(&L, 0, sizeof(L));  // padding included: the bytes are the graph's key
        L.np = e.d_npairs.as<const int>();
        L.nj = e.d_npairs.as<const int>() + cB;
        L.cp = e.d_cp.as<const rlnk::ChunkPtrs>();
        L.cap_pairs = e.need_pairs;
        L.cap_jobs = e.need_jobs;
        L.cap_small = e.need_small;
        L.cap_wide = e.need_wide;
        L.grid_B = cB;
        L.pair_off = e.d_pair_off.as<long long>();
        L.fine_n = e.d_fine_n.as<long long>();
        L.foff = e.d_foff.as<long long>();
        L.job_off = e.d_job_off.as<int>();
        L.pairs = e.d_pairs.as<int>();
        L.plan = e.d_plan.as<rlnk::FinePlan>();
        L.mask = e.d_mask.as<const unsigned char>();
        L.d2off = d_d2off.as<const long long>();
        L.or_cnt = e.d_or_cnt.as<const int>();
        L.oc_off = e.d_oc_off.as<const int>();
        L.ocnt = e.d_ocnt.as<const int>();
        L.NT = NT;
        L.jobs = e.d_jobs.as<rlnk::FineJob>();
        L.job_order = e.d_job_order.as<int>();
        L.or_off = e.d_or_off.as<const int>();
        L.or_frame = e.d_or_frame.as<const int>();
        L.or_class = e.d_or_class.as<const int>();
        L.n_or_grid = e.grid.n_dir * e.grid.n_psi;
        {
            const long long n_keys = (long long)K * L.n_or_grid;
            int shift = 0;
            while (((n_keys - 1) >> shift) >= rlnk::kJobKeys) ++shift;
            L.key_shift = shift;
        }
        L.sorted = kFineOrders[e.fine_order].sorted;
        L.items_s = e.d_items_s.as<int2>();
        L.items_w = e.d_items_w.as<int2>();
        rlnk::FineArgs &fa = L.fa;
        fa.jobs = e.d_jobs.as<const rlnk::FineJob>();
        fa.items_s = e.d_items_s.as<const int2>();
        fa.items_w = e.d_items_w.as<const int2>();
        fa.plan = e.d_plan.as<rlnk::FinePlan>();
        fa.cp = e.d_cp.as<const rlnk::ChunkPtrs>();
        fa.pairs = e.d_pairs.as<const int>();
        fa.pair_off = e.d_pair_off.as<const long long>();
        fa.or_off = e.d_or_off.as<const int>();
        fa.or_frame = e.d_or_frame.as<const int>();
        fa.or_class = e.d_or_class.as<const int>();
        fa.vols = e.d_vols.as<const float4 *const>();
        fa.vg = e.vg;
        fa.fine_frames = e.d_frames_f.as<const rlnk::Frame>();
        fa.ig = e.igf;
        fa.scale = d_scale;
        fa.q = e.d_q_f.as<const float>();
        fa.ex = e.d_ex_f.as<const float2>();
        fa.ey = e.d_ey_f.as<const float2>();
        fa.tr_ix = e.d_tr_ix.as<const int>();
        fa.tr_iy = e.d_tr_iy.as<const int>();
        fa.pix = e.d_band_pix.as<const int2>() + (size_t)e.band_choice * e.pf_sampled;
        fa.n_pix = (int)e.pf_sampled;
        fa.n_bands = kBandChoices[e.band_choice];
        fa.item_major = kFineOrders[e.fine_order].item_major;
        fa.xcd_queues = kFineOrders[e.fine_order].xcd_queues;
        L.fbase = e.d_fbase.as<const double>();
        L.shape = e.fine_shape;
        fa.NT = NT;
        fa.do_ctf = cfg.do_ctf_correction;
        fa.diff2 = e.d_fdiff2.as<double>();
        rlnk::WeightArgs &wf = L.wf;
        wf = wa;
        wf.diff2 = nullptr;
        wf.diff2d = e.d_fdiff2.as<const double>();
        wf.d2_off = e.d_foff.as<const long long>();
        wf.n_entries = e.d_fine_n.as<const long long>();
        wf.pairs = e.d_pairs.as<const int>();
        wf.pair_off = e.d_pair_off.as<const long long>();
        wf.fine = 1;
        wf.maxsig = 0;
        wf.weight = e.d_fweight.as<double>();
        wf.mask = nullptr;
        wf.stats = e.d_stats_f.as<rlnk::PassStats>();
        wf.plan = e.d_plan.as<const rlnk::FinePlan>();
        L.shell = e.d_shell_acc.as<double>();
        L.part = e.d_part_acc.as<double>();
        L.pdfc = e.d_pdfc_acc.as<double>();
        L.n_shell = (long long)e.nshell * cB;
        L.n_part = 8LL * cB;
        L.n_pdfc = (long long)K * cB;
        rlnk::WsumArgs &ws = L.ws;
        ws.jobs = e.d_jobs.as<const rlnk::FineJob>();
        ws.pairs = e.d_pairs.as<const int>();
        ws.pair_off = e.d_pair_off.as<const long long>();
        ws.or_off = e.d_or_off.as<const int>();
        ws.or_frame = e.d_or_frame.as<const int>();
        ws.or_class = e.d_or_class.as<const int>();
        ws.or_dir = e.d_or_dir.as<const int>();
        ws.vols = e.d_vols.as<const float4 *const>();
        ws.vg = e.vg;
        ws.fine_frames = e.d_frames_f.as<const rlnk::Frame>();
        ws.ig = e.igf;
        ws.cp = e.d_cp.as<const rlnk::ChunkPtrs>();
        ws.plan = e.d_plan.as<rlnk::FinePlan>();
        ws.scale_w = d_scale_w;
        ws.minv_w = e.d_minv_w.as<const float>();
        ws.mresol = e.d_mresol.as<const int>();
        ws.wpix = e.d_wsum_pix.as<const int2>() + (size_t)wsum_band_choice(e) * e.n_wsum_pix;
        ws.n_wpix = e.n_wsum_pix;
        ws.n_bands = kBandChoices[wsum_band_choice(e)];
        ws.unit_order = e.wsum_order;
        ws.job_order = e.d_job_order.as<const int>();
        ws.xa_shell = (cfg.do_scale_correction && e.d_xa_shell.ptr) ? e.d_xa_shell.as<const unsigned char>() : nullptr;
        ws.ex = e.d_ex_f.as<const float2>();
        ws.ey = e.d_ey_f.as<const float2>();
        ws.tr_ix = e.d_tr_ix.as<const int>();
        ws.tr_iy = e.d_tr_iy.as<const int>();
        ws.over_x = e.d_over_x.as<const double>();
        ws.over_y = e.d_over_y.as<const double>();
        ws.weight = e.d_fweight.as<const double>();
        ws.stats = e.d_stats_f.as<const rlnk::PassStats>();
        ws.old_off = d_old;
        ws.prior_off = d_prior;
        ws.pixel_size = cfg.pixel_size;
        ws.NT = NT;
        ws.nshell = e.nshell;
        ws.n_dir = e.grid.n_dir;
        ws.do_ctf = cfg.do_ctf_correction;
        ws.do_scale = cfg.do_scale_correction;
        ws.bp_slots = e.d_bp_ptrs.as<float *const>();
        ws.shell_acc = e.d_shell_acc.as<double>();
        ws.part_acc = e.d_part_acc.as<double>();
        ws.pdf_class = e.d_pdfc_acc.as<double>();
        ws.pdf_direction = e.d_pdfd_acc.as<double>();
        ws.K = K;
        L.grid_fine = e.grid_fine;
        L.grid_wsum = e.grid_wsum;
        RLN_TRY(run_fine_stage(e, L));
        HIP_TRY(hipMemcpyAsync(&plan, e.d_plan.ptr, sizeof(plan), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (!plan.overflow) break;
        if (attempt >= 2) RLN_FAIL(RLN_ERR_MEMORY, "fine pass capacity not reached after %d attempts", attempt + 1);
        // grow to what the plan asked for, with headroom, and run the pass again
        e.need_pairs = std::max(e.need_pairs, plan.n_pairs + plan.n_pairs / 4 + 1024);
        e.need_jobs = std::max(e.need_jobs, plan.n_jobs + plan.n_jobs / 4 + 256);
        e.need_small = std::max(e.need_small, plan.n_small + plan.n_small / 4 + 256);
        e.need_wide = std::max(e.need_wide, plan.n_wide + plan.n_wide / 4 + 256);
        e.counters[13] += 1;
    }
    const long long n_pairs = plan.n_pairs;
    const int n_jobs = plan.n_jobs;
    std::vector<int> counts(B);
    std::vector<long long> pair_off(B), foff(B);
    std::vector<rlnk::PassStats> stc(B);
    HIP_TRY(hipMemcpyAsync(counts.data(), e.d_npairs.ptr, sizeof(int) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(pair_off.data(), e.d_pair_off.ptr, sizeof(long long) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(stc.data(), e.d_stats_c.ptr, sizeof(rlnk::PassStats) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < B; ++b) foff[b] = 32LL * pair_off[b];
    for (int b = 0; b < B; ++b)
        if (stc[b].status != 0) {
            int code = stc[b].status;
            RLN_FAIL(code, "coarse pass of particle %d: %s", b,
                     code == RLN_ERR_SUM_WEIGHT_ZERO ? "zero or NaN sum of weights" : "no significant samples");
        }
    HIP_TRY(hipEventRecord(e.ev[kStCount], st));
    // results
    std::vector<rlnk::PassStats> stf(B);
    std::vector<double> shells(e.nshell), pacc(8 * (size_t)B), pdfc((size_t)K * B);
    hipLaunchKernelGGL(rlnk::k_shell_total, dim3(e.nshell), dim3(rlnk::kBlock), 0, st, e.d_shell_acc.as<const double>(),
                       B, e.nshell, e.d_shell_tot.as<double>());
    RLN_TRY(launch_check("shell sums"));
    HIP_TRY(hipMemcpyAsync(stf.data(), e.d_stats_f.ptr, sizeof(rlnk::PassStats) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(shells.data(), e.d_shell_tot.ptr, sizeof(double) * shells.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(pacc.data(), e.d_part_acc.ptr, sizeof(double) * pacc.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(pdfc.data(), e.d_pdfc_acc.ptr, sizeof(double) * pdfc.size(), hipMemcpyDeviceToHost, st));
    // best fine pose per particle: the pair's coarse index and the job's orientation entry
    std::vector<int> all_pairs(std::max<long long>(1, n_pairs));
    if (n_pairs)
        HIP_TRY(hipMemcpyAsync(all_pairs.data(), e.d_pairs.ptr, sizeof(int) * n_pairs, hipMemcpyDeviceToHost, st));
    if (e.keep) {
        e.kept.assign(B, Engine::Kept());
        for (int b = 0; b < B; ++b) {
            Engine::Kept &kp = e.kept[b];
            kp.n_coarse = h.n_entries[b];
            kp.n_dir = h.nd[b];
            kp.n_psi = h.np[b];
            kp.cdiff2.resize(kp.n_coarse);
            kp.cweight.resize(kp.n_coarse);
            kp.cmask.resize(kp.n_coarse);
            HIP_TRY(hipMemcpyAsync(kp.cdiff2.data(), e.d_diff2.as<float>() + h.d2_off[b], sizeof(float) * kp.n_coarse,
                                   hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(kp.cweight.data(), e.d_weight.as<double>() + h.d2_off[b],
                                   sizeof(double) * kp.n_coarse, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(kp.cmask.data(), e.d_mask.as<unsigned char>() + h.d2_off[b], kp.n_coarse,
                                   hipMemcpyDeviceToHost, st));
            const long long nf = 32LL * counts[b];
            std::vector<int> pr(counts[b]);
            kp.fdiff2.resize(nf);
            kp.fweight.resize(nf);
            kp.f_index.resize(nf);
            if (counts[b]) {
                HIP_TRY(hipMemcpyAsync(pr.data(), e.d_pairs.as<int>() + pair_off[b], sizeof(int) * counts[b],
                                       hipMemcpyDeviceToHost, st));
                HIP_TRY(hipMemcpyAsync(kp.fdiff2.data(), e.d_fdiff2.as<double>() + foff[b], sizeof(double) * nf,
                                       hipMemcpyDeviceToHost, st));
                HIP_TRY(hipMemcpyAsync(kp.fweight.data(), e.d_fweight.as<double>() + foff[b], sizeof(double) * nf,
                                       hipMemcpyDeviceToHost, st));
            }
            HIP_TRY(hipStreamSynchronize(st));
            for (long long f = 0; f < nf; ++f) kp.f_index[f] = (long long)pr[f >> 5] * 32 + (f & 31);
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    const clk::time_point t_epi = clk::now();
    for (int i = 0; i < kStCount; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.ev[i], e.ev[i + 1]) != hipSuccess) ms = 0.f;
        e.stage_ms[i] += ms;
    }
    (void)hipGetLastError();  // a stage without its events leaves no sticky error for the next launch check
    e.counters[0] += B;
    e.counters[1] += h.total_entries;
    e.counters[2] = (long long)Pc;
    e.counters[3] += 32 * n_pairs;
    e.counters[4] = (long long)Pf;
    e.counters[5] += (long long)plan.n_small + plan.n_wide;
    for (int b = 0; b < B; ++b) {
        e.counters[6] += (long long)pacc[8 * b + 5];
        e.counters[7] += (long long)pacc[8 * b + 6];
    }
    e.counters[9] += (long long)h.tiles.size();
    e.counters[10] += n_jobs;
    e.counters[11] += 1;
    e.counters[14] = e.grid_fine;
    e.counters[15] = e.grid_wsum;
    e.counters[16] = e.pf_sampled;
    e.counters[17] = kBandChoices[e.band_choice];
    e.counters[18] = rlnk::kFineShapes[e.fine_shape].tc;
    e.counters[19] = rlnk::kFineShapes[e.fine_shape].cw;
    e.counters[20] = e.fine_order;
    e.counters[21] = e.wsum_order;
    e.counters[22] = kBandChoices[wsum_band_choice(e)];
    for (int b = 0; b < B; ++b) {
        if (stf[b].status != 0) RLN_FAIL(stf[b].status, "fine pass of particle %d: zero or NaN sum of weights", b);
        if (stf[b].best_index < 0) RLN_FAIL(RLN_ERR_FILTERED_ZERO, "no positive fine weight for particle %d", b);
    }
    // per-particle results (the best pose decoded to angles and offsets) on several host threads; the sums over
    // particles below in particle order
    auto decode = [&](int b0, int b1) {
        for (int b = b0; b < b1; ++b) {
            const rlnk::PassStats &sf = stf[b];
            rln_particle_result &r = results[b];
            std::memset(&r, 0, sizeof(r));
            r.coarse_min_diff2 = stc[b].min_diff2;
            r.coarse_sum_weight = stc[b].sum_weight;
            r.coarse_significant_weight = stc[b].threshold;
            r.nr_significant = (double)stc[b].nr_sign;
            r.min_diff2 = sf.min_diff2;
            r.sum_weight = sf.sum_weight;
            r.significant_weight = sf.threshold;
            r.nr_fine_evaluated = 32 * counts[b];
            r.pmax = sf.max_weight / sf.sum_weight;
            // decode the best pose
            const int c = all_pairs[pair_off[b] + (sf.best_index >> 5)];
            const int o = c / NT, t = c % NT;
            const int rr = (int)((sf.best_index >> 2) & 7), u = (int)(sf.best_index & 3);
            const int og = h.or_frame[h.or_off[b] + o] / 8, k = h.or_class[h.or_off[b] + o];
            e.grid.oversampled_orientation(og / e.grid.n_psi, og % e.grid.n_psi, rr, r.rot, r.tilt, r.psi);
            r.xoff = meta[b].old_offset_x + e.over_x[t * 4 + u];
            r.yoff = meta[b].old_offset_y + e.over_y[t * 4 + u];
            r.iclass = k + 1;
            r.best_ihidden_over = ((long long)o * NT + t) * 32 + rr * 4 + u;
            r.wsum_norm_correction = pacc[8 * b + 2];
            double xa = pacc[8 * b + 3], aa = pacc[8 * b + 4];
            if (cfg.do_scale_correction) {
                xa /= meta[b].scale;
                aa /= meta[b].scale * meta[b].scale;
            }
            r.signal_product = xa;
            r.reference_power = aa;
            r.dLL = log(sf.sum_weight) - sf.min_diff2 - e.logsigma2;
        }
    };
    {
        const int nt = B >= 1024 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t)
            pool.emplace_back(decode, (int)((long long)B * t / nt), (int)((long long)B * (t + 1) / nt));
        decode(0, (int)((long long)B / nt));
        for (auto &th : pool) th.join();
    }
    if (wsum && wsum->sigma2_noise)
        for (int s = 0; s < e.nshell; ++s) wsum->sigma2_noise[s] += shells[s];
    for (int b = 0; b < B; ++b) {
        const rln_particle_result &r = results[b];
        const double xa = r.signal_product, aa = r.reference_power;
        if (wsum) {
            if (wsum->pdf_class)
                for (int kk = 0; kk < K; ++kk) wsum->pdf_class[kk] += pdfc[(size_t)b * K + kk];
            wsum->sumw_group += pacc[8 * b + 0];
            wsum->sigma2_offset += pacc[8 * b + 1];
            wsum->signal_product += xa;
            wsum->reference_power += aa;
            wsum->LL += r.dLL;
            wsum->ave_Pmax += r.pmax;
        }
    }
    e.host_ms[2] += ms_since(t_epi);
    return RLN_OK;
}

}  // namespace rln

struct rln_estep : rln::Engine {};

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
extern "C" const char *rln_last_error(void) { return rln::last_error(); }

extern "C" int rln_projector_create(int device, rln_projector **out) {
    if (!out) RLN_FAIL(RLN_ERR_ARG, "rln_projector_create: null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) RLN_FAIL(RLN_ERR_ARG, "rln_projector_create: no HIP device %d", device);
    rln_projector *p = new rln_projector();
    p->device = device;
    *out = p;
    return RLN_OK;
}

extern "C" int rln_projector_set_mdl_dim(rln_projector *p, int xdim, int ydim, int zdim, int inity, int initz,
                                         int maxr, double padding_factor, int *changed) {
    if (!p) RLN_FAIL(RLN_ERR_ARG, "null projector");
    rln::MdlDims d{xdim, ydim, zdim, inity, initz, maxr, padding_factor};
    RLN_TRY(rln::check_dims(d));
    const bool same = d == p->dims && p->data.ptr;
    if (changed) *changed = same ? 0 : 1;
    if (same) return RLN_OK;
    HIP_TRY(hipSetDevice(p->device));
    p->data.release();
    p->dims = d;
    RLN_TRY(p->data.ensure(4 * sizeof(float4) * p->bases()));
    p->version++;
    return RLN_OK;
}

extern "C" int rln_projector_init_mdl(rln_projector *p, const float *real, const float *imag) {
    if (!p || !p->data.ptr || !real || !imag) RLN_FAIL(RLN_ERR_STATE, "projector not dimensioned or null data");
    const size_t n = p->dims.voxels();
    std::vector<float2> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = make_float2(real[i], imag[i]);
    HIP_TRY(hipSetDevice(p->device));
    RLN_TRY(rln::pack_into_projector(p, tmp.data()));
    p->version++;
    return RLN_OK;
}

extern "C" int rln_projector_init_mdl_complex(rln_projector *p, const double *data) {
    if (!p || !p->data.ptr || !data) RLN_FAIL(RLN_ERR_STATE, "projector not dimensioned or null data");
    const size_t n = p->dims.voxels();
    std::vector<float2> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = make_float2((float)data[2 * i], (float)data[2 * i + 1]);
    HIP_TRY(hipSetDevice(p->device));
    RLN_TRY(rln::pack_into_projector(p, tmp.data()));
    p->version++;
    return RLN_OK;
}

extern "C" int rln_projector_destroy(rln_projector *p) {
    if (!p) return RLN_OK;
    (void)hipSetDevice(p->device);
    p->data.release();
    delete p;
    return RLN_OK;
}

extern "C" int rln_backprojector_create(int device, rln_backprojector **out) {
    if (!out) RLN_FAIL(RLN_ERR_ARG, "rln_backprojector_create: null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) RLN_FAIL(RLN_ERR_ARG, "rln_backprojector_create: no HIP device %d", device);
    rln_backprojector *b = new rln_backprojector();
    b->device = device;
    *out = b;
    return RLN_OK;
}

extern "C" int rln_backprojector_set_mdl_dim(rln_backprojector *b, int xdim, int ydim, int zdim, int inity, int initz,
                                             int maxr, double padding_factor, size_t *bytes) {
    if (!b) RLN_FAIL(RLN_ERR_ARG, "null backprojector");
    rln::MdlDims d{xdim, ydim, zdim, inity, initz, maxr, padding_factor};
    RLN_TRY(rln::check_dims(d));
    HIP_TRY(hipSetDevice(b->device));
    if (!(d == b->dims) || !b->slots.ptr) {
        b->slots.release();
        b->dims = d;
        RLN_TRY(b->slots.ensure(sizeof(float) * rlnk::kSlot * b->bases()));
    }
    if (bytes) *bytes = sizeof(float) * rlnk::kSlot * b->bases();
    return RLN_OK;
}

extern "C" int rln_backprojector_init_mdl(rln_backprojector *b) {
    if (!b || !b->slots.ptr) RLN_FAIL(RLN_ERR_STATE, "backprojector not dimensioned");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipMemset(b->slots.ptr, 0, sizeof(float) * rlnk::kSlot * b->bases()));
    return RLN_OK;
}

extern "C" int rln_backprojector_get_mdl_data(rln_backprojector *b, float *real, float *imag, float *weight) {
    if (!b || !b->slots.ptr) RLN_FAIL(RLN_ERR_STATE, "backprojector not dimensioned");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipDeviceSynchronize());
    const size_t nv = b->dims.voxels();
    rln::DevBuf plain;
    RLN_TRY(plain.ensure(3 * sizeof(float) * nv));
    float *d_re = plain.as<float>(), *d_im = d_re + nv, *d_w = d_im + nv;
    const rlnk::VolGeom g{b->dims.xdim, b->dims.ydim, b->dims.zdim, b->dims.inity, b->dims.initz};
    hipLaunchKernelGGL(rlnk::k_fold_slots, dim3(rlnk::fold_blocks(g)), dim3(rlnk::kBlock), 0, 0,
                       b->slots.as<const float>(), g, 0, d_re, d_im, d_w);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (real) HIP_TRY(hipMemcpy(real, d_re, sizeof(float) * nv, hipMemcpyDeviceToHost));
    if (imag) HIP_TRY(hipMemcpy(imag, d_im, sizeof(float) * nv, hipMemcpyDeviceToHost));
    if (weight) HIP_TRY(hipMemcpy(weight, d_w, sizeof(float) * nv, hipMemcpyDeviceToHost));
    return RLN_OK;
}

// Staging of rln_backprojector_get_mdl_data_noalloc: brick layers of the folded data, imaginary and weight volumes.
// Each device holds its own copy until the process ends. Host code passes the device address that
// hipGetSymbolAddress returns for the current device (kept per device below), never the symbol itself.
static constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kScratchDevices] = {};

// Device address of this device's scratch array; looked up once per device, with that device current.
int scratch_on_device(int device, float **out) {
    if (device < 0 || device >= kScratchDevices) RLN_FAIL(RLN_ERR_ARG, "device %d beyond %d", device, kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) RLN_FAIL(RLN_ERR_HIP, "no device address for the staging array on device %d", device);
        g_scratch_ptr[device] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[device];
    return RLN_OK;
}
}  // namespace

extern "C" int rln_backprojector_get_mdl_data_noalloc(rln_backprojector *b, float *real, float *imag, float *weight) {
    if (!b || !b->slots.ptr) RLN_FAIL(RLN_ERR_STATE, "backprojector not dimensioned");
    const rln::MdlDims &d = b->dims;
    const size_t plane = (size_t)d.xdim * d.ydim;
    const size_t layer_bytes = 3 * sizeof(float) * 4 * plane;  // one brick layer (4 planes) of re, im, weight
    if (layer_bytes > kNoAllocScratchBytes)
        RLN_FAIL(RLN_ERR_ARG, "rln_backprojector_get_mdl_data_noalloc: 4 planes of %d x %d voxels need %zu bytes of "
                 "staging, beyond its %zu (xdim * ydim <= %zu)", d.xdim, d.ydim, layer_bytes, kNoAllocScratchBytes,
                 kNoAllocScratchBytes / (12 * sizeof(float)));
    HIP_TRY(hipSetDevice(b->device));
    float *sp = nullptr;
    RLN_TRY(scratch_on_device(b->device, &sp));
    HIP_TRY(hipDeviceSynchronize());
    const rlnk::VolGeom g{d.xdim, d.ydim, d.zdim, d.inity, d.initz};
    const int n_layers = (d.zdim + 3) / 4;
    const int per_pass = (int)(kNoAllocScratchBytes / layer_bytes);
    for (int l0 = 0; l0 < n_layers; l0 += per_pass) {
        const int nl = std::min(per_pass, n_layers - l0);
        const int z0 = 4 * l0, nz = std::min(4 * nl, d.zdim - z0);
        const size_t stride = 4 * (size_t)nl * plane;  // floats per component in the staging array
        float *d_re = sp, *d_im = sp + stride, *d_w = sp + 2 * stride;
        hipLaunchKernelGGL(rlnk::k_fold_slots, dim3(rlnk::fold_blocks(g, nl)), dim3(rlnk::kBlock), 0, 0,
                           b->slots.as<const float>(), g, l0, d_re, d_im, d_w);
        HIP_TRY(hipGetLastError());
        // synchronous copies on the null stream: the next pass overwrites the staging only after they complete
        const size_t off = (size_t)z0 * plane, bytes = sizeof(float) * (size_t)nz * plane;
        if (real) HIP_TRY(hipMemcpy(real + off, d_re, bytes, hipMemcpyDeviceToHost));
        if (imag) HIP_TRY(hipMemcpy(imag + off, d_im, bytes, hipMemcpyDeviceToHost));
        if (weight) HIP_TRY(hipMemcpy(weight + off, d_w, bytes, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipDeviceSynchronize());
    return RLN_OK;
}

extern "C" int rln_backprojector_destroy(rln_backprojector *b) {
    if (!b) return RLN_OK;
    (void)hipSetDevice(b->device);
    b->slots.release();
    delete b;
    return RLN_OK;
}

extern "C" int rln_estep_create(int device, const rln_estep_config *cfg, rln_estep **out) {
    if (!cfg || !out) RLN_FAIL(RLN_ERR_ARG, "rln_estep_create: null argument");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) RLN_FAIL(RLN_ERR_ARG, "rln_estep_create: no HIP device %d", device);
    rln_estep *e = new rln_estep();
    e->device = device;
    const int rc = rln::engine_init(*e, *cfg);
    if (rc != RLN_OK) {
        std::string msg = rln::last_error();
        rln_estep_destroy(e);
        rln::set_error("%s", msg.c_str());
        return rc;
    }
    *out = e;
    return RLN_OK;
}

extern "C" int rln_estep_destroy(rln_estep *e) {
    if (!e) return RLN_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    rln::DevBuf *bufs[] = {&e->d_frames_c, &e->d_frames_f, &e->d_proj_c, &e->d_pf, &e->d_ehi, &e->d_elo,
                           &e->d_zero_row, &e->d_wc, &e->d_grp, &e->d_exg, &e->d_ctf2, &e->d_ynorm, &e->d_q_c, &e->d_q_f, &e->d_minv_w,
                           &e->d_mresol, &e->d_xa_shell, &e->d_ex_c, &e->d_ey_c, &e->d_ex_f, &e->d_ey_f,
                           &e->d_tr_ix, &e->d_tr_iy, &e->d_over_x, &e->d_over_y, &e->d_vols, &e->d_bp_ptrs,
                           &e->d_fimg, &e->d_fimgn, &e->d_fctf, &e->d_meta_f, &e->d_meta_d, &e->d_meta_ll,
                           &e->d_or_off, &e->d_oc_off, &e->d_or_cnt, &e->d_or_proj, &e->d_or_prior, &e->d_or_frame,
                           &e->d_or_class, &e->d_or_dir, &e->d_tr_prior, &e->d_ctfq, &e->d_yq, &e->d_tiles,
                           &e->d_diff2, &e->d_weight, &e->d_mask, &e->d_stats_c, &e->d_stats_f, &e->d_npairs,
                           &e->d_njobs, &e->d_nchunks, &e->d_pair_off, &e->d_job_off, &e->d_chunk_off, &e->d_pairs,
                           &e->d_jobs, &e->d_job_order, &e->d_chunk_job, &e->d_chunk_first, &e->d_fdiff2, &e->d_fbase, &e->d_fweight,
                           &e->d_fine_n, &e->d_nent, &e->d_foff, &e->d_shell_acc, &e->d_part_acc, &e->d_pdfc_acc, &e->d_pdfd_acc,
                           &e->d_shell_tot, &e->d_plan, &e->d_cp, &e->d_items_s, &e->d_items_w, &e->d_wsegs, &e->d_seg_first,
                           &e->d_wpart, &e->d_hcnt, &e->d_hsum, &e->d_wcand, &e->d_segbest, &e->d_ocnt};
    rln::drop_fine_graphs(*e);
    for (rln::DevBuf *b : bufs) b->release();
    for (int i = 0; i <= rln::kStCount; ++i)
        if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return RLN_OK;
}

extern "C" int rln_estep_set_model(rln_estep *e, rln_projector *const *projectors,
                                   rln_backprojector *const *backprojectors, const double *pdf_class,
                                   const double *pdf_direction, const double *sigma2_noise,
                                   const double *data_vs_prior) {
    if (!e || !projectors || !backprojectors || !pdf_class || !sigma2_noise)
        RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_model: null argument");
    if (!e->cfg.local_search && !pdf_direction) RLN_FAIL(RLN_ERR_ARG, "pdf_direction required for global searches");
    std::lock_guard<std::mutex> lk(e->mtx);
    const int K = e->cfg.nr_classes;
    const int r_max = std::min(e->cfg.current_size / 2, e->cfg.ori_size / 2);
    const int pad_size = 2 * (rln::ROUND(e->cfg.padding_factor * r_max) + 1) + 1;
    for (int k = 0; k < K; ++k) {
        const rln::MdlDims &d = projectors[k]->dims;
        if (!projectors[k]->data.ptr || d.xdim != pad_size / 2 + 1 || d.ydim != pad_size || d.zdim != pad_size ||
            d.inity != -(pad_size / 2) || d.initz != -(pad_size / 2))
            RLN_FAIL(RLN_ERR_ARG, "projector %d dims do not match current_size %d, padding %g", k, e->cfg.current_size,
                     e->cfg.padding_factor);
        const rln::MdlDims &bd = backprojectors[k]->dims;
        if (!backprojectors[k]->slots.ptr || !(bd == d))
            RLN_FAIL(RLN_ERR_ARG, "backprojector %d dims do not match the projector", k);
        if (projectors[k]->device != e->device || backprojectors[k]->device != e->device)
            RLN_FAIL(RLN_ERR_ARG, "projector/backprojector %d on another device", k);
    }
    e->K = K;
    e->proj.assign(projectors, projectors + K);
    e->bproj.assign(backprojectors, backprojectors + K);
    e->proj_version.assign(K, ~0ULL);
    e->pdf_class.assign(pdf_class, pdf_class + K);
    const int nd = e->grid.n_dir;
    if (pdf_direction)
        e->pdf_direction.assign(pdf_direction, pdf_direction + (size_t)K * nd);
    else
        e->pdf_direction.assign((size_t)K * nd, 1.);
    // mean of the global-search orientation priors, summed in the order of the exhaustive enumeration (class,
    // direction, psi) so that explicit orientation lists reproduce its normalisation bit for bit
    double pm = 0.;
    for (int k = 0; k < K; ++k)
        for (int d = 0; d < nd; ++d)
            for (int ip = 0; ip < e->grid.n_psi; ++ip) pm += e->pdf_direction[(size_t)k * nd + d];
    pm /= (double)((long long)K * nd * e->grid.n_psi);
    e->pdf_dir_mean = pm != 0. ? pm : 1.;
    e->sigma2_noise.assign(sigma2_noise, sigma2_noise + e->nshell);
    if (data_vs_prior)
        e->data_vs_prior.assign(data_vs_prior, data_vs_prior + (size_t)K * e->nshell);
    else
        e->data_vs_prior.clear();
    const rln::MdlDims &d = projectors[0]->dims;
    e->vg = rlnk::VolGeom{d.xdim, d.ydim, d.zdim, d.inity, d.initz};
    HIP_TRY(hipSetDevice(e->device));
    RLN_TRY(rln::engine_model_tables(*e));
    e->coarse_valid = false;
    e->model_set = true;
    return RLN_OK;
}

extern "C" int rln_estep_set_pool(rln_estep *e, int max_particles) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_pool: null engine");
    if (max_particles < 1) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_pool: %d particles per chunk", max_particles);
    std::lock_guard<std::mutex> lk(e->mtx);
    e->max_pool = max_particles;
    return RLN_OK;
}

extern "C" int rln_estep_use_graphs(rln_estep *e, int enable) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_use_graphs: null engine");
    std::lock_guard<std::mutex> lk(e->mtx);
    e->graph_off = !enable;
    if (!enable) {
        rln::drop_fine_graphs(*e);
        e->fine_key.clear();
    }
    return RLN_OK;
}

extern "C" int rln_estep_fine_tuning(rln_estep *e, double *ms, int capacity, int *n_out) {
    if (!e || !n_out) RLN_FAIL(RLN_ERR_ARG, "rln_estep_fine_tuning: null argument");
    std::lock_guard<std::mutex> lk(e->mtx);
    const int n = (int)e->tune_ms.size();
    *n_out = n;
    if (ms)
        for (int i = 0; i < std::min(n, capacity); ++i) ms[i] = e->tune_ms[i];
    return RLN_OK;
}

extern "C" int rln_estep_set_fine_bands(rln_estep *e, int n_bands) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_bands: null engine");
    std::lock_guard<std::mutex> lk(e->mtx);
    if (n_bands == 0) {
        e->tuned = false;
        return RLN_OK;
    }
    for (int c = 0; c < rln::kNumBandChoices; ++c)
        if (rln::kBandChoices[c] == n_bands) {
            e->band_choice = c;
            e->grid_fine = e->resident_fine > 0 ? e->resident_fine : 1024;
            e->tuned = true;
            return RLN_OK;
        }
    RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_bands: %d bands (1, 2, 4, 8, 16, 32 or 0)", n_bands);
}

extern "C" int rln_estep_set_fine_order(rln_estep *e, int order, int wsum_order) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_order: null engine");
    if (order < -1 || order >= rln::kNumFineOrders || wsum_order < 0 || wsum_order > 2)
        RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_order: order %d (-1 .. %d), weighted-sum order %d (0, 1 or 2)", order,
                 rln::kNumFineOrders - 1, wsum_order);
    std::lock_guard<std::mutex> lk(e->mtx);
    e->order_fixed = order;
    if (order >= 0) e->fine_order = order;
    e->wsum_order = wsum_order;
    return RLN_OK;
}

extern "C" int rln_estep_set_fine_shape(rln_estep *e, int shape) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_shape: null engine");
    if (shape < -1 || shape >= rlnk::kNumFineShapes)
        RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_fine_shape: shape %d (-1 .. %d)", shape, rlnk::kNumFineShapes - 1);
    std::lock_guard<std::mutex> lk(e->mtx);
    e->shape_fixed = shape;
    if (shape >= 0) e->fine_shape = shape;
    return RLN_OK;
}

extern "C" int rln_estep_set_wsum_bands(rln_estep *e, int n_bands) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_wsum_bands: null engine");
    bool ok = n_bands == 0;
    for (int c = 0; c < rln::kNumBandChoices; ++c) ok = ok || rln::kBandChoices[c] == n_bands;
    if (!ok) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_wsum_bands: %d bands (1, 2, 4, 8, 16, 32 or 0)", n_bands);
    std::lock_guard<std::mutex> lk(e->mtx);
    e->wsum_bands = n_bands;
    return RLN_OK;
}

extern "C" int rln_estep_set_orientation_sets(rln_estep *e, int n, const int *set_offsets, const int *set_orients) {
    if (!e || n < 0 || (n > 0 && !set_offsets)) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_orientation_sets: null argument");
    std::lock_guard<std::mutex> lk(e->mtx);
    if (e->cfg.local_search) RLN_FAIL(RLN_ERR_ARG, "orientation lists replace the global search; local_search is set");
    const int n_or = e->grid.n_dir * e->grid.n_psi;
    if (set_offsets[0] != 0) RLN_FAIL(RLN_ERR_ARG, "set_offsets[0] must be 0");
    for (int i = 0; i < n; ++i)
        if (set_offsets[i + 1] < set_offsets[i]) RLN_FAIL(RLN_ERR_ARG, "set_offsets not ascending at %d", i);
    const int total = set_offsets[n];
    if (total > 0 && !set_orients) RLN_FAIL(RLN_ERR_ARG, "rln_estep_set_orientation_sets: null orientation list");
    for (int j = 0; j < total; ++j)
        if (set_orients[j] < 0 || set_orients[j] >= n_or)
            RLN_FAIL(RLN_ERR_ARG, "orientation %d out of range [0, %d)", set_orients[j], n_or);
    e->set_off.assign(set_offsets, set_offsets + n + 1);
    e->set_og.assign(set_orients, set_orients + total);
    return RLN_OK;
}

extern "C" int rln_estep_run(rln_estep *e, int n, const float *fimg, const float *fimg_nomask, const float *fctf,
                             int inputs_on_device, const rln_particle_meta *meta, rln_particle_result *results,
                             rln_wsum *wsum) {
    if (!e || n < 0 || (n > 0 && (!fimg || !fctf || !meta || !results)))
        RLN_FAIL(RLN_ERR_ARG, "rln_estep_run: null argument");
    std::lock_guard<std::mutex> lk(e->mtx);
    if (!e->model_set) RLN_FAIL(RLN_ERR_STATE, "rln_estep_run: model not set");
    if (!fimg_nomask) fimg_nomask = fimg;
    HIP_TRY(hipSetDevice(e->device));
    for (int i = 0; i < rln::kStCount; ++i) e->stage_ms[i] = 0.;
    for (double &h : e->host_ms) h = 0.;
    for (long long &c : e->counters) c = 0;
    const auto t_run = std::chrono::steady_clock::now();
    HIP_TRY(hipMemsetAsync(e->d_pdfd_acc.ptr, 0, sizeof(double) * (size_t)e->K * e->grid.n_dir, e->stream));
    RLN_TRY(rln::engine_coarse_projections(*e));
    // explicit orientation lists apply to this run only
    std::vector<int> set_off, set_og;
    set_off.swap(e->set_off);
    set_og.swap(e->set_og);
    const bool have_sets = !set_off.empty();
    if (have_sets && (int)set_off.size() != n + 1)
        RLN_FAIL(RLN_ERR_ARG, "orientation lists were set for %d particles, run has %d", (int)set_off.size() - 1, n);
    // chunk boundaries from a device-memory budget for the coarse intermediates (13 bytes per coarse pose:
    // diff2, FP64 weight, mask), with per-particle pose counts: the explicit list, the whole grid, or a fraction
    // of it for local searches
    const size_t Pc = e->igc.P, Pf = e->igf.P;
    const bool split = e->cfg.precision == RLN_PREC_SPLIT_BF16;
    const double fixed = (split ? 6.0 * e->Kp : 8.0 * Pc * e->NT) + 20.0 * Pf + 64.0 * 1024;
    const double budget = 24.0e9;
    std::vector<int> starts;
    {
        // the first chunk holds at most a quarter of the pool and at most 1,024 particles: its lists are built
        // before any device work of the run, the later chunks' while the device runs the chunk before
        const int first_pool = std::max(1, std::min(e->max_pool / 4, 1024));
        double used = 0.;
        int cnt = 0;
        for (int i = 0; i < n; ++i) {
            long long orients = e->cfg.local_search ? std::max(1, e->n_or / 8) : e->n_or;
            if (have_sets && set_off[i + 1] > set_off[i]) orients = set_off[i + 1] - set_off[i];
            const double need = 13.0 * (double)e->K * (double)orients * e->NT + fixed;
            const int pool = starts.size() == 1 ? first_pool : e->max_pool;
            if (cnt == 0 || (!e->keep && (used + need > budget || cnt >= pool))) {
                starts.push_back(i);
                used = 0.;
                cnt = 0;
            }
            used += need;
            ++cnt;
        }
        starts.push_back(n);
    }
    for (size_t c = 0; c + 1 < starts.size(); ++c) e->cap_B = std::max(e->cap_B, starts[c + 1] - starts[c]);
    // chunk c runs on the device while a second host thread builds the lists of chunk c + 1
    using clk = std::chrono::steady_clock;
    for (auto &ch : e->chunk_host)
        if (!ch) ch.reset(new rln::ChunkHost());
    rln::ChunkHost *pcur = e->chunk_host[0].get(), *pnxt = e->chunk_host[1].get();
    auto so = [&](int i) { return have_sets ? set_off.data() + i : nullptr; };
    auto sg = [&](int i) { return have_sets ? set_og.data() + set_off[i] : nullptr; };
    const int n_chunks = (int)starts.size() - 1;
    if (n_chunks > 0) {
        const clk::time_point t0 = clk::now();
        rln::build_chunk_host(*e, starts[1] - starts[0], meta, *pcur, so(0), sg(0));
        e->host_ms[1] = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        e->host_ms[0] += e->host_ms[1];
    }
    for (int c = 0; c < n_chunks; ++c) {
        const int i0 = starts[c], B = starts[c + 1] - starts[c];
        bool next_ok = true;
        std::thread builder;
        if (c + 1 < n_chunks) {
            const int i1 = starts[c + 1], B1 = starts[c + 2] - starts[c + 1];
            // chunk c + 1 packs into the staging buffer of chunk c - 1, whose copies ended with that chunk
            builder = std::thread([&, i1, B1, c] {
                try {
                    rln::build_chunk_host(*e, B1, meta + i1, *pnxt, so(i1), sg(i1));
                    if (rln::pack_chunk_host(*e, *pnxt, e->h_stage[(c + 1) & 1]) != RLN_OK) next_ok = false;
                } catch (...) {
                    next_ok = false;
                }
            });
        }
        const size_t off_c = (size_t)i0 * Pf;
        const int rc = rln::run_chunk(*e, B, fimg + 2 * off_c, fimg_nomask + 2 * off_c, fctf + off_c,
                                      inputs_on_device, meta + i0, *pcur, results + i0, wsum);
        if (builder.joinable()) {
            const clk::time_point t0 = clk::now();
            builder.join();
            e->host_ms[0] += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        }
        if (rc != RLN_OK) return rc;
        if (!next_ok) RLN_FAIL(RLN_ERR_MEMORY, "rln_estep_run: building the lists of a chunk failed");
        std::swap(pcur, pnxt);
    }
    e->have_timing = true;
    if (wsum && wsum->pdf_direction) {
        std::vector<double> pd((size_t)e->K * e->grid.n_dir);
        HIP_TRY(hipMemcpyAsync(pd.data(), e->d_pdfd_acc.ptr, sizeof(double) * pd.size(), hipMemcpyDeviceToHost,
                               e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (size_t i = 0; i < pd.size(); ++i) wsum->pdf_direction[i] += pd[i];
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->host_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_run).count();
    return RLN_OK;
}

extern "C" int rln_selftest_mfma_bf16(int device, const float *A, const float *B, float *D) {
    if (!A || !B || !D) RLN_FAIL(RLN_ERR_ARG, "null argument");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) RLN_FAIL(RLN_ERR_ARG, "no HIP device %d", device);
    HIP_TRY(hipSetDevice(device));
    rln::DevBuf buf;
    RLN_TRY(buf.ensure(sizeof(float) * (16 * 32 + 32 * 16 + 16 * 16)));
    float *dA = buf.as<float>(), *dB = dA + 16 * 32, *dD = dB + 32 * 16;
    HIP_TRY(hipMemcpy(dA, A, sizeof(float) * 16 * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dB, B, sizeof(float) * 32 * 16, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(rlnk::k_selftest_mfma, dim3(1), dim3(64), 0, 0, dA, dB, dD);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(D, dD, sizeof(float) * 16 * 16, hipMemcpyDeviceToHost));
    return RLN_OK;
}

extern "C" int rln_estep_keep_intermediates(rln_estep *e, int keep) {
    if (!e) RLN_FAIL(RLN_ERR_ARG, "null engine");
    e->keep = keep != 0;
    if (!e->keep) e->kept.clear();
    return RLN_OK;
}

extern "C" int rln_estep_coarse_size(rln_estep *e, int ip, long long *n_coarse, int *n_dir, int *n_psi,
                                     int *n_trans) {
    if (!e || ip < 0 || ip >= (int)e->kept.size()) RLN_FAIL(RLN_ERR_STATE, "no kept intermediates for particle %d", ip);
    const rln::Engine::Kept &k = e->kept[ip];
    if (n_coarse) *n_coarse = k.n_coarse;
    if (n_dir) *n_dir = k.n_dir;
    if (n_psi) *n_psi = k.n_psi;
    if (n_trans) *n_trans = e->NT;
    return RLN_OK;
}

extern "C" int rln_estep_get_coarse(rln_estep *e, int ip, float *diff2, double *weight, unsigned char *significant) {
    if (!e || ip < 0 || ip >= (int)e->kept.size()) RLN_FAIL(RLN_ERR_STATE, "no kept intermediates for particle %d", ip);
    const rln::Engine::Kept &k = e->kept[ip];
    if (diff2) std::copy(k.cdiff2.begin(), k.cdiff2.end(), diff2);
    if (weight) std::copy(k.cweight.begin(), k.cweight.end(), weight);
    if (significant) std::copy(k.cmask.begin(), k.cmask.end(), significant);
    return RLN_OK;
}

extern "C" int rln_estep_fine_size(rln_estep *e, int ip, long long *n_fine) {
    if (!e || ip < 0 || ip >= (int)e->kept.size()) RLN_FAIL(RLN_ERR_STATE, "no kept intermediates for particle %d", ip);
    *n_fine = (long long)e->kept[ip].fdiff2.size();
    return RLN_OK;
}

extern "C" int rln_estep_get_fine(rln_estep *e, int ip, long long *ihidden_over, double *diff2, double *weight) {
    if (!e || ip < 0 || ip >= (int)e->kept.size()) RLN_FAIL(RLN_ERR_STATE, "no kept intermediates for particle %d", ip);
    const rln::Engine::Kept &k = e->kept[ip];
    if (ihidden_over) std::copy(k.f_index.begin(), k.f_index.end(), ihidden_over);
    if (diff2) std::copy(k.fdiff2.begin(), k.fdiff2.end(), diff2);
    if (weight) std::copy(k.fweight.begin(), k.fweight.end(), weight);
    return RLN_OK;
}

extern "C" int rln_estep_last_counters(rln_estep *e, long long *out, int n_max, int *n_out) {
    if (!e || !out) RLN_FAIL(RLN_ERR_ARG, "null argument");
    const int n = std::min(n_max, (int)(sizeof(e->counters) / sizeof(e->counters[0])));
    for (int i = 0; i < n; ++i) out[i] = e->counters[i];
    if (n_out) *n_out = n;
    return RLN_OK;
}

extern "C" int rln_estep_last_timings(rln_estep *e, double *ms, int n_max, int *n_out) {
    if (!e || !ms) RLN_FAIL(RLN_ERR_ARG, "null argument");
    const int n = std::min<int>(n_max, rln::kStCount + 4);
    for (int i = 0; i < n; ++i) ms[i] = i < rln::kStCount ? e->stage_ms[i] : e->host_ms[i - rln::kStCount];
    if (n_out) *n_out = n;
    return RLN_OK;
}
