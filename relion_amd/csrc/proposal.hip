// Learned pose proposal behind include/rln_proposal.h: training-set generation from the reference, a six-layer
// convolutional network trained with hand-written forward and backward kernels (proposal_kernels.h), and top-k
// bin prediction that seeds the E-step's coarse orientation lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../include/rln_proposal.h"
#include "host_common.h"
#include "proposal_kernels.h"
#include "sampling.h"

namespace rln {
namespace {

constexpr int kLayers = 6;
constexpr int kStride[kLayers] = {1, 2, 1, 2, 1, 2};

struct ConvLayer {
    int cin, cout, H, Ho, stride, Kp;
    size_t w_off, b_off;
};

unsigned grid1(long long n) { return (unsigned)((n + rlnp::kB - 1) / rlnp::kB); }

}  // namespace

}  // namespace rln

struct rln_proposal {
    int device = 0;
    int n_cu = 0;  // compute units of the device (persistent grids)
    // grid orientations of every bin for the last (grid order, grid psi count) asked, built once
    mutable std::mutex orient_mtx;
    mutable std::vector<std::vector<int>> orient_cache;
    mutable int orient_order = -1, orient_npsi = -1;
    rln_proposal_config cfg{};
    hipStream_t stream = nullptr;
    // backward pass: the weight gradients' partial-sum reductions run on `side` beside the data gradients;
    // ev_dw orders each after its weight-gradient kernel, ev_red the next weight-gradient kernel (and the end of the
    // pass) after the last reduction
    hipStream_t side = nullptr;
    hipEvent_t ev_dw = nullptr, ev_red = nullptr;
    bool red_pending = false;
    std::mutex mtx;
    int W = 0, C = 0, NB = 0, F = 0, n_dir_bins = 0;
    rln::ConvLayer conv[rln::kLayers];
    size_t fcw_off = 0, fcb_off = 0, n_params = 0;
    rln::DevBuf w, g, m1, m2, wb;  // FP32 master, gradient, Adam moments; BF16 copy
    int adam_t = 0;
    std::mt19937_64 rng;
    // activations for a batch of up to cap images
    int cap = 0;
    rln::DevBuf x[rln::kLayers + 1], col[rln::kLayers], logits, dlog, dtmp, dz_a, dz_b, loss, labels, idx;
    rln::DevBuf wt;  // the layers' weights rearranged for their data gradients, layer l at l * wt_slot
    size_t wt_slot = 0;
    rln::DevBuf dzt;  // a 64-channel layer's output gradient transposed to [64][M] for its weight gradient
    rln::DevBuf part;  // partial sums of a k-split GEMM
    rln::DevBuf dwpart;  // per-block weight and bias gradient sums of k_conv_dw_patch
    rln::DevBuf colpart;  // per-block column sums of k_conv_bwd_s2_patch
    rln::DevBuf src, src_ctf;  // particle images and CTFs staged from the host
    rln::DevBuf top_bins, top_probs;  // predictions of one predict call
    // a prediction launched and not yet finished: its size, the pinned host copies of its top-k, its end event
    int pend_n = -1, pend_topk = 0;
    void *h_top = nullptr;  // pinned: int bins[n][topk] then float probs[n][topk]
    size_t h_top_bytes = 0;
    hipEvent_t pend_ev = nullptr;
    // training set
    rln::DevBuf train_img;  // bf16 [n_train][W][W]
    std::vector<int> train_lab;
    long long n_train = 0, train_cap = 0;
};

namespace rln {
namespace {

// The data gradient of a layer runs as the implicit GEMM k_conv_bwd_data when its channel counts fit it.
bool bwd_data_direct(const ConvLayer &L) { return L.cin % 16 == 0 && L.cin <= 64 && L.cout % 8 == 0; }

// Bias gradients: out[j] += column sums of Z [M][N] (bf16)
void col_sum(rln_proposal &p, const rlnp::bf16 *Z, int M, int N, float *out);
bool dw_patch_geom(const ConvLayer &L, rlnp::PatchGeom &g, size_t &lds);

int ensure_batch(rln_proposal &p, int n) {
    if (n <= p.cap) return RLN_OK;
    // the convolution kernels index a layer's n x H x H positions in 32 bits
    if ((long long)n * p.W * p.W >= (1LL << 31))
        RLN_FAIL(RLN_ERR_ARG, "batch of %d images of %d x %d pixels: at most 2^31 pixels per layer", n, p.W, p.W);
    const int cap = std::max(n, 64);
    const int W = p.W, C = p.C;
    RLN_TRY(p.x[0].ensure(sizeof(rlnp::bf16) * (size_t)cap * W * W));
    for (int l = 0; l < kLayers; ++l) {
        const ConvLayer &L = p.conv[l];
        const size_t M = (size_t)cap * L.Ho * L.Ho;
        rlnp::PatchGeom dg{};
        size_t dlds = 0;
        if (!dw_patch_geom(L, dg, dlds)) RLN_TRY(p.col[l].ensure(sizeof(rlnp::bf16) * M * L.Kp));
        RLN_TRY(p.x[l + 1].ensure(sizeof(rlnp::bf16) * M * C));
    }
    size_t dmax = (size_t)cap * std::max(p.NB, p.F);
    for (int l = 1; l < kLayers; ++l)  // dcol of the layers whose data gradient takes the col2im path
        if (!bwd_data_direct(p.conv[l]))
            dmax = std::max(dmax, (size_t)cap * p.conv[l].Ho * p.conv[l].Ho * p.conv[l].Kp);
    size_t wtmax = 1;
    for (int l = 1; l < kLayers; ++l) wtmax = std::max(wtmax, (size_t)9 * p.conv[l].cin * p.conv[l].cout);
    p.wt_slot = wtmax;
    RLN_TRY(p.wt.ensure(sizeof(rlnp::bf16) * wtmax * kLayers));  // one slot per layer
    RLN_TRY(p.logits.ensure(sizeof(float) * (size_t)cap * p.NB));
    RLN_TRY(p.dlog.ensure(sizeof(rlnp::bf16) * (size_t)cap * p.NB));
    RLN_TRY(p.dtmp.ensure(sizeof(float) * dmax));
    const size_t zmax = (size_t)cap * W * W * C;
    RLN_TRY(p.dz_a.ensure(sizeof(rlnp::bf16) * zmax));
    RLN_TRY(p.dz_b.ensure(sizeof(rlnp::bf16) * zmax));
    RLN_TRY(p.dzt.ensure(sizeof(rlnp::bf16) * zmax));
    RLN_TRY(p.dwpart.ensure(sizeof(float) * rlnp::kDwPart * (size_t)(p.n_cu > 0 ? p.n_cu : 256)));
    RLN_TRY(p.colpart.ensure(sizeof(float) * 64 * (size_t)(p.n_cu > 0 ? p.n_cu : 256)));
    RLN_TRY(p.loss.ensure(sizeof(double) * cap));
    RLN_TRY(p.labels.ensure(sizeof(int) * cap));
    RLN_TRY(p.idx.ensure(sizeof(int) * cap));
    p.cap = cap;
    return RLN_OK;
}

template <bool AK, bool BK, int EPI>
int gemm(rln_proposal &p, const rlnp::bf16 *A, long long sam, long long sak, const rlnp::bf16 *B, long long sbk,
         long long sbn, int M, int N, int K, const float *bias, float *Cf, rlnp::bf16 *Cb, long long ldc,
         int kchunk = 0) {
    const long long tiles = (long long)((M + 63) / 64) * ((N + 63) / 64);
    if (EPI == rlnp::kEpiF32 && kchunk <= 0) {
        // few output tiles (the fully connected layer of a small batch): k split into nz chunks of at least 512 so
        // that the launch holds about four blocks per CU, partial sums per chunk added in chunk order
        const long long cu = p.n_cu > 0 ? p.n_cu : 256;
        int nz = (int)std::min<long long>({8, (4 * cu + tiles - 1) / tiles, K / 512});
        if (nz > 1) {
            const int kc = 32 * ((K + 32 * nz - 1) / (32 * nz));
            nz = (K + kc - 1) / kc;
            RLN_TRY(p.part.ensure(sizeof(float) * (size_t)nz * M * ldc));
            rlnp::GemmArgs a{A, B, sam, sak, sbk, sbn, M, N, K, kc, bias, p.part.as<float>(), nullptr, ldc,
                             (long long)M * ldc};
            hipLaunchKernelGGL((rlnp::k_gemm<AK, BK, rlnp::kEpiF32>), dim3((M + 63) / 64, (N + 63) / 64, nz),
                               dim3(rlnp::kB), 0, p.stream, a);
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(rlnp::k_sum_parts, dim3(grid1((long long)M * N)), dim3(rlnp::kB), 0, p.stream,
                               p.part.as<const float>(), nz, M, N, ldc, Cf);
            HIP_TRY(hipGetLastError());
            return RLN_OK;
        }
    }
    if (kchunk <= 0) kchunk = 32 * ((K + 31) / 32);
    rlnp::GemmArgs a{A, B, sam, sak, sbk, sbn, M, N, K, kchunk, bias, Cf, Cb, ldc, 0};
    dim3 grid((M + 63) / 64, (N + 63) / 64, (K + kchunk - 1) / kchunk);
    if (EPI == rlnp::kEpiAtomicF32 && grid.z == 1)  // one writer per element: a plain read-add-write
        hipLaunchKernelGGL((rlnp::k_gemm<AK, BK, rlnp::kEpiAddF32>), grid, dim3(rlnp::kB), 0, p.stream, a);
    else
        hipLaunchKernelGGL((rlnp::k_gemm<AK, BK, EPI>), grid, dim3(rlnp::kB), 0, p.stream, a);
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

const rlnp::bf16 *wbf(const rln_proposal &p, size_t off) { return p.wb.as<const rlnp::bf16>() + off; }

template <int NT>
void conv_fwd_t(rln_proposal &p, const ConvLayer &L, int n, rlnp::bf16 *col = nullptr) {
    const long long M = (long long)n * L.Ho * L.Ho;
    hipLaunchKernelGGL((rlnp::k_conv_fwd<NT>), dim3((unsigned)((M + 63) / 64)), dim3(rlnp::kB), 0, p.stream,
                       p.x[&L - p.conv].as<const rlnp::bf16>(), n, L.H, L.cin, L.stride, L.Ho, L.Kp,
                       p.wb.as<const rlnp::bf16>() + L.w_off, p.w.as<const float>() + L.b_off,
                       p.x[&L - p.conv + 1].as<rlnp::bf16>(), col);
}

// LDS-resident weights (Cin % 8 == 0, Kp % 32 == 0, weights within 80 KB so that two blocks share a CU)
bool conv_lds_ok(const ConvLayer &L) {
    return L.cin % 8 == 0 && L.Kp % 32 == 0 && (size_t)L.cout * (L.Kp + rlnp::kConvLdsPad) * 2 <= 80 * 1024;
}

template <int NT>
int conv_fwd_lds_t(rln_proposal &p, const ConvLayer &L, int n) {
    const long long M = (long long)n * L.Ho * L.Ho;
    const size_t lds = (size_t)16 * NT * (L.Kp + rlnp::kConvLdsPad) * sizeof(rlnp::bf16);
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&rlnp::k_conv_fwd_lds<NT>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (M + rlnp::kConvTile - 1) / rlnp::kConvTile;
    const unsigned grid = (unsigned)std::min<long long>(tiles, 1024);
    hipLaunchKernelGGL((rlnp::k_conv_fwd_lds<NT>), dim3(grid), dim3(rlnp::kConvThreads), lds, p.stream,
                       p.x[&L - p.conv].as<const rlnp::bf16>(), n, L.H, L.cin, L.stride, L.Ho, L.Kp,
                       p.wb.as<const rlnp::bf16>() + L.w_off, p.w.as<const float>() + L.b_off,
                       p.x[&L - p.conv + 1].as<rlnp::bf16>());
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

// The 64-channel layers with staged input patches (k_conv_fwd_patch): tiles of 256 / stride output pixels that start
// at row or image boundaries, and a patch that fits beside the weights in a CU's 160 KB of LDS.
bool patch_geom(const ConvLayer &L, rlnp::PatchGeom &g, size_t &lds) {
    if (L.cin != 64 || L.cout != 64 || L.Kp != rlnp::kPatchK || (L.stride != 1 && L.stride != 2)) return false;
    g.H = L.H;
    g.Ho = L.Ho;
    g.TP = 256 / L.stride;
    const int hoho = L.Ho * L.Ho;
    int th;
    if (g.TP >= hoho) {
        if (g.TP % hoho) return false;
        g.n_img_t = g.TP / hoho;
        th = L.Ho;
    } else {
        if (hoho % g.TP || g.TP % L.Ho) return false;
        g.n_img_t = 1;
        th = g.TP / L.Ho;
    }
    g.PR = (th - 1) * L.stride + 3;
    g.units = g.n_img_t * g.PR * L.H * 8;
    if (g.units > rlnp::kPatchRegs * rlnp::kPatchThreads) return false;
    lds = sizeof(rlnp::bf16) * ((size_t)64 * (rlnp::kPatchK + rlnp::kConvLdsPad) +
                                (size_t)g.n_img_t * g.PR * L.H * rlnp::kPatchPix);
    return lds <= 160 * 1024;
}

// The stride-2 64-channel layers' data gradient (k_conv_bwd_s2_patch): tiles of 256 positions of the dZ grid that
// start at row or image boundaries, each staged with the dZ row after it.
bool bwd_s2_geom(const ConvLayer &L, rlnp::PatchGeom &g, size_t &lds) {
    if (L.cin != 64 || L.cout != 64 || L.Kp != rlnp::kPatchK || L.stride != 2 || L.H != 2 * L.Ho) return false;
    g.H = L.H;
    g.Ho = L.Ho;
    g.TP = 256;
    const int hoho = L.Ho * L.Ho;
    int th;
    if (g.TP >= hoho) {
        if (g.TP % hoho) return false;
        g.n_img_t = g.TP / hoho;
        th = L.Ho;
    } else {
        if (hoho % g.TP || g.TP % L.Ho) return false;
        g.n_img_t = 1;
        th = g.TP / L.Ho;
    }
    g.PR = th + 1;
    g.units = g.n_img_t * g.PR * L.Ho * 8;
    if (g.units > rlnp::kPatchRegs * rlnp::kPatchThreads) return false;
    // weights, the dZ patch and the 8 waves' output staging (32 positions each)
    lds = sizeof(rlnp::bf16) * ((size_t)64 * (rlnp::kPatchK + rlnp::kConvLdsPad) +
                                ((size_t)g.n_img_t * g.PR * L.Ho + 8 * 32) * rlnp::kPatchPix);
    return lds <= 160 * 1024;
}

// A first layer (one input channel, 9 taps padded to 16) runs its forward pass as k_conv_fwd in training too, which
// stores the im2col columns of its weight gradient on the way.
bool direct_first(const ConvLayer &L) { return L.cin == 1 && L.Kp == 16; }

// The 64-channel layers whose weight and bias gradients k_conv_dw_patch computes from the staged input rows (the
// forward's tiles; output rows of whole 8-position fragments, H = stride * Ho); its LDS bytes in lds.
bool dw_patch_geom(const ConvLayer &L, rlnp::PatchGeom &g, size_t &lds) {
    size_t flds = 0;
    if (!patch_geom(L, g, flds) || L.Ho % 8 || L.H != L.stride * L.Ho || g.TP != 256 / L.stride) return false;
    if ((g.n_img_t * g.PR * L.H + 63) / 64 > rlnp::dw_xregs(L.stride)) return false;  // staged pixels: 64 per register
    lds = sizeof(rlnp::bf16) * rlnp::dw_lds_elems(g, L.stride);
    return lds <= 160 * 1024;
}

// k_conv_fwd_patch over n images of geometry g: input X, weights Wt, bias (EPI 0) or ReLU mask Ymask (EPI 1), output Y
template <int S, int EPI>
int conv_patch_launch(rln_proposal &p, int n, const rlnp::PatchGeom &g, size_t lds, const rlnp::bf16 *X,
                      const rlnp::bf16 *Wt, const float *bias, rlnp::bf16 *Y, const rlnp::bf16 *Ymask) {
    const long long M = (long long)n * g.Ho * g.Ho;
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&rlnp::k_conv_fwd_patch<S, EPI>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (M + g.TP - 1) / g.TP;
    const unsigned grid = (unsigned)std::min<long long>(tiles, p.n_cu > 0 ? p.n_cu : 256);
    hipLaunchKernelGGL((rlnp::k_conv_fwd_patch<S, EPI>), dim3(grid), dim3(rlnp::kPatchThreads), lds, p.stream, X, n,
                       g, Wt, bias, Y, Ymask);
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

template <int S>
int conv_fwd_patch_t(rln_proposal &p, const ConvLayer &L, int n, const rlnp::PatchGeom &g, size_t lds) {
    return conv_patch_launch<S, 0>(p, n, g, lds, p.x[&L - p.conv].as<const rlnp::bf16>(),
                                   p.wb.as<const rlnp::bf16>() + L.w_off, p.w.as<const float>() + L.b_off,
                                   p.x[&L - p.conv + 1].as<rlnp::bf16>(), nullptr);
}

// logits of the n images in x[0]; with keep_col the im2col buffers are written (the weight gradients of the
// backward pass read them), otherwise the convolutions run as implicit GEMMs straight from the activations. The
// 64-channel layers compute their outputs with k_conv_fwd_patch either way, so training and prediction round alike.
int forward(rln_proposal &p, int n, bool keep_col = true) {
    for (int l = 0; l < kLayers; ++l) {
        const ConvLayer &L = p.conv[l];
        const long long M = (long long)n * L.Ho * L.Ho;
        rlnp::PatchGeom pg{};
        size_t plds = 0;
        if (patch_geom(L, pg, plds)) {
            rlnp::PatchGeom dg{};
            size_t dlds = 0;
            if (keep_col && !dw_patch_geom(L, dg, dlds)) {  // k-major columns (colT[k][m]) for the weight gradient
                hipLaunchKernelGGL(rlnp::k_im2col_t64, dim3((unsigned)((M + 63) / 64), 9), dim3(rlnp::kB), 0,
                                   p.stream, p.x[l].as<const rlnp::bf16>(), n, L.H, L.stride, L.Ho,
                                   p.col[l].as<rlnp::bf16>());
                HIP_TRY(hipGetLastError());
            }
            if (L.stride == 1) RLN_TRY(conv_fwd_patch_t<1>(p, L, n, pg, plds));
            else RLN_TRY(conv_fwd_patch_t<2>(p, L, n, pg, plds));
            continue;
        }
        if (!keep_col && L.cout % 16 == 0 && L.cout <= 64 && conv_lds_ok(L)) {
            switch (L.cout / 16) {
                case 1: RLN_TRY(conv_fwd_lds_t<1>(p, L, n)); break;
                case 2: RLN_TRY(conv_fwd_lds_t<2>(p, L, n)); break;
                case 3: RLN_TRY(conv_fwd_lds_t<3>(p, L, n)); break;
                default: RLN_TRY(conv_fwd_lds_t<4>(p, L, n)); break;
            }
            continue;
        }
        if ((!keep_col || direct_first(L)) && L.cout % 16 == 0 && L.cout <= 64) {
            rlnp::bf16 *col = keep_col ? p.col[l].as<rlnp::bf16>() : nullptr;
            switch (L.cout / 16) {
                case 1: conv_fwd_t<1>(p, L, n, col); break;
                case 2: conv_fwd_t<2>(p, L, n, col); break;
                case 3: conv_fwd_t<3>(p, L, n, col); break;
                default: conv_fwd_t<4>(p, L, n, col); break;
            }
            HIP_TRY(hipGetLastError());
            continue;
        }
        const long long threads = (L.cin % 8 == 0) ? M * 9 * (L.cin / 8) : M * L.Kp;
        hipLaunchKernelGGL(rlnp::k_im2col, dim3(grid1(threads)), dim3(rlnp::kB), 0, p.stream,
                           p.x[l].as<const rlnp::bf16>(), n, L.H, L.cin, L.stride, L.Ho, L.Kp, p.col[l].as<rlnp::bf16>());
        HIP_TRY(hipGetLastError());
        RLN_TRY((gemm<true, true, rlnp::kEpiBiasReluBf16>(
            p, p.col[l].as<const rlnp::bf16>(), L.Kp, 1, wbf(p, L.w_off), 1, L.Kp, (int)M, L.cout, L.Kp,
            p.w.as<const float>() + L.b_off, nullptr, p.x[l + 1].as<rlnp::bf16>(), L.cout)));
    }
    RLN_TRY((gemm<true, true, rlnp::kEpiF32>(p, p.x[kLayers].as<const rlnp::bf16>(), p.F, 1, wbf(p, p.fcw_off), 1,
                                              p.F, n, p.NB, p.F, p.w.as<const float>() + p.fcb_off,
                                              p.logits.as<float>(), nullptr, p.NB)));
    return RLN_OK;
}

// gradients of the mean cross entropy into p.g (accumulated); per-image losses into p.loss
int backward(rln_proposal &p, int n) {
    const float scale = 1.f / (float)n;
    hipLaunchKernelGGL(rlnp::k_softmax_xent, dim3(n), dim3(rlnp::kB), 0, p.stream, p.logits.as<const float>(),
                       p.labels.as<const int>(), p.NB, scale, p.loss.as<double>(), p.dlog.as<rlnp::bf16>());
    HIP_TRY(hipGetLastError());
    float *gw = p.g.as<float>();
    const rlnp::bf16 *x6 = p.x[kLayers].as<const rlnp::bf16>();
    const rlnp::bf16 *dl = p.dlog.as<const rlnp::bf16>();
    // fully connected layer
    RLN_TRY((gemm<false, false, rlnp::kEpiAtomicF32>(p, dl, 1, p.NB, x6, p.F, 1, p.NB, p.F, n, nullptr,
                                                      gw + p.fcw_off, nullptr, p.F)));
    col_sum(p, dl, n, p.NB, gw + p.fcb_off);
    HIP_TRY(hipGetLastError());
    RLN_TRY((gemm<true, false, rlnp::kEpiF32>(p, dl, p.NB, 1, wbf(p, p.fcw_off), p.F, 1, n, p.F, p.NB, nullptr,
                                               p.dtmp.as<float>(), nullptr, p.F)));
    rlnp::bf16 *dz = p.dz_a.as<rlnp::bf16>(), *dz_next = p.dz_b.as<rlnp::bf16>();
    hipLaunchKernelGGL(rlnp::k_mask_relu, dim3(grid1((long long)n * p.F)), dim3(rlnp::kB), 0, p.stream,
                       p.dtmp.as<const float>(), x6, (long long)n * p.F, dz);
    HIP_TRY(hipGetLastError());
    bool bias_below_done = false;  // this layer's bias gradient came with the layer above's data gradient
    {  // the 64-channel layers' rearranged weights for their data gradients, in one launch
        rlnp::WtPlan plan{};
        int ne = 0;
        for (int l = 1; l < kLayers; ++l) {
            rlnp::PatchGeom g{};
            size_t lds = 0;
            const ConvLayer &L = p.conv[l];
            const int mode = (L.stride == 1 && patch_geom(L, g, lds)) ? 1 : bwd_s2_geom(L, g, lds) ? 0 : -1;
            if (mode < 0) continue;
            plan.w_off[ne] = (long long)L.w_off;
            plan.off_out[ne] = (long long)(l * p.wt_slot);
            plan.mode[ne] = mode;
            ++ne;
        }
        if (ne > 0) {
            hipLaunchKernelGGL(rlnp::k_wt_layers, dim3(grid1(64 * rlnp::kPatchK), ne), dim3(rlnp::kB), 0, p.stream,
                               p.wb.as<const rlnp::bf16>(), plan, p.wt.as<rlnp::bf16>());
            HIP_TRY(hipGetLastError());
        }
    }
    for (int l = kLayers - 1; l >= 0; --l) {
        const ConvLayer &L = p.conv[l];
        const int M = n * L.Ho * L.Ho;
        // dW = dZ^T col, split over the rows (the pixels)
        const int kchunk = 32 * std::max(1, std::min((M + 31) / 32, 64));
        rlnp::PatchGeom pg{};
        size_t plds = 0;
        bool bias_done = false;
        if (dw_patch_geom(L, pg, plds)) {
            // weight and bias gradients from the staged input rows and dZ transposed, per-block sums reduced
            hipLaunchKernelGGL(rlnp::k_transpose64, dim3((unsigned)((M + 63) / 64)), dim3(rlnp::kB), 0, p.stream, dz,
                               (long long)M, p.dzt.as<rlnp::bf16>());
            HIP_TRY(hipGetLastError());
            const long long tiles = ((long long)M + pg.TP - 1) / pg.TP;
            const int nb = (int)std::min<long long>(tiles, p.n_cu > 0 ? p.n_cu : 256);
            if ((size_t)nb * rlnp::kDwPart * sizeof(float) > p.dwpart.bytes)
                RLN_FAIL(RLN_ERR_STATE, "weight-gradient partial sums: %d blocks exceed the buffer", nb);
            if (p.red_pending) HIP_TRY(hipStreamWaitEvent(p.stream, p.ev_red, 0));  // the partials were read
            if (L.stride == 1) {
                HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&rlnp::k_conv_dw_patch<1>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds));
                hipLaunchKernelGGL(rlnp::k_conv_dw_patch<1>, dim3(nb), dim3(rlnp::kPatchThreads), plds, p.stream,
                                   p.x[l].as<const rlnp::bf16>(), n, pg, p.dzt.as<const rlnp::bf16>(),
                                   p.dwpart.as<float>());
            } else {
                HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&rlnp::k_conv_dw_patch<2>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds));
                hipLaunchKernelGGL(rlnp::k_conv_dw_patch<2>, dim3(nb), dim3(rlnp::kPatchThreads), plds, p.stream,
                                   p.x[l].as<const rlnp::bf16>(), n, pg, p.dzt.as<const rlnp::bf16>(),
                                   p.dwpart.as<float>());
            }
            HIP_TRY(hipGetLastError());
            // the bias follows the weights in the parameter vector: one reduction covers both; it runs on the side
            // stream while this layer's data gradient runs
            HIP_TRY(hipEventRecord(p.ev_dw, p.stream));
            HIP_TRY(hipStreamWaitEvent(p.side, p.ev_dw, 0));
            hipLaunchKernelGGL(rlnp::k_dw_reduce, dim3((unsigned)((rlnp::kDwPart / 4 + rlnp::kB - 1) / rlnp::kB),
                                                        (unsigned)std::max(1, nb / 16)),
                               dim3(rlnp::kB), 0, p.side, p.dwpart.as<const float>(), nb, gw + L.w_off);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(p.ev_red, p.side));
            p.red_pending = true;
            bias_done = true;
        } else if (patch_geom(L, pg, plds)) {
            // k-major columns (forward) and dZ transposed: both operands with the pixel index contiguous
            hipLaunchKernelGGL(rlnp::k_transpose64, dim3((unsigned)((M + 63) / 64)), dim3(rlnp::kB), 0, p.stream, dz,
                               (long long)M, p.dzt.as<rlnp::bf16>());
            HIP_TRY(hipGetLastError());
            RLN_TRY((gemm<true, true, rlnp::kEpiAtomicF32>(p, p.dzt.as<const rlnp::bf16>(), M, 1,
                                                            p.col[l].as<const rlnp::bf16>(), 1, M, L.cout, L.Kp, M,
                                                            nullptr, gw + L.w_off, nullptr, L.Kp, kchunk)));
        } else {
            RLN_TRY((gemm<false, false, rlnp::kEpiAtomicF32>(p, dz, 1, L.cout, p.col[l].as<const rlnp::bf16>(), L.Kp,
                                                              1, L.cout, L.Kp, M, nullptr, gw + L.w_off, nullptr,
                                                              L.Kp, kchunk)));
        }
        if (!bias_done && !bias_below_done) col_sum(p, dz, M, L.cout, gw + L.b_off);
        bias_below_done = false;
        HIP_TRY(hipGetLastError());
        if (l == 0) break;
        if (L.stride == 1 && patch_geom(L, pg, plds)) {
            // a stride-1 64-channel layer: its data gradient is a stride-1 convolution of dZ with the flipped,
            // transposed weights, masked by the ReLU of its input (same geometry as its forward pass)
            RLN_TRY((conv_patch_launch<1, 1>(p, n, pg, plds, dz, p.wt.as<const rlnp::bf16>() + l * p.wt_slot, nullptr,
                                             dz_next, p.x[l].as<const rlnp::bf16>())));
        } else if (bwd_s2_geom(L, pg, plds)) {
            // a stride-2 64-channel layer: the four parity classes of each 2 x 2 input block from a staged dZ patch
            HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&rlnp::k_conv_bwd_s2_patch),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds));
            const long long tiles = ((long long)n * L.Ho * L.Ho + pg.TP - 1) / pg.TP;
            const unsigned grid = (unsigned)std::min<long long>(tiles, p.n_cu > 0 ? p.n_cu : 256);
            // the layer below's bias gradient from the masked dX, unless its own weight-gradient kernel sums it
            rlnp::PatchGeom dg{};
            size_t dlds = 0;
            const bool colsum = !dw_patch_geom(p.conv[l - 1], dg, dlds) &&
                                (size_t)grid * 64 * sizeof(float) <= p.colpart.bytes;
            hipLaunchKernelGGL(rlnp::k_conv_bwd_s2_patch, dim3(grid), dim3(rlnp::kPatchThreads), plds, p.stream, dz, n,
                               pg, p.wt.as<const rlnp::bf16>() + l * p.wt_slot, p.x[l].as<const rlnp::bf16>(),
                               dz_next, colsum ? p.colpart.as<float>() : nullptr);
            HIP_TRY(hipGetLastError());
            if (colsum) {
                hipLaunchKernelGGL(rlnp::k_colsum_parts, dim3(1), dim3(1024), 0, p.stream, p.colpart.as<const float>(),
                                   (int)grid, gw + p.conv[l - 1].b_off);
                HIP_TRY(hipGetLastError());
                bias_below_done = true;
            }
        } else if (bwd_data_direct(L)) {
            const int nwt = 9 * L.cin * L.cout;
            hipLaunchKernelGGL(rlnp::k_wt_bwd, dim3(grid1(nwt)), dim3(rlnp::kB), 0, p.stream, wbf(p, L.w_off), L.cin,
                               L.cout, L.Kp, p.wt.as<rlnp::bf16>());
            const long long rows = (long long)n * L.H * L.H;
            const dim3 grid((unsigned)((rows + 63) / 64));
            const rlnp::bf16 *wd = p.wt.as<const rlnp::bf16>(), *y = p.x[l].as<const rlnp::bf16>();
            switch (L.cin / 16) {
                case 1: hipLaunchKernelGGL(rlnp::k_conv_bwd_data<1>, grid, dim3(rlnp::kB), 0, p.stream, dz, n, L.H, L.cout, L.stride, L.Ho, wd, y, dz_next); break;
                case 2: hipLaunchKernelGGL(rlnp::k_conv_bwd_data<2>, grid, dim3(rlnp::kB), 0, p.stream, dz, n, L.H, L.cout, L.stride, L.Ho, wd, y, dz_next); break;
                case 3: hipLaunchKernelGGL(rlnp::k_conv_bwd_data<3>, grid, dim3(rlnp::kB), 0, p.stream, dz, n, L.H, L.cout, L.stride, L.Ho, wd, y, dz_next); break;
                default: hipLaunchKernelGGL(rlnp::k_conv_bwd_data<4>, grid, dim3(rlnp::kB), 0, p.stream, dz, n, L.H, L.cout, L.stride, L.Ho, wd, y, dz_next); break;
            }
            HIP_TRY(hipGetLastError());
        } else {
            RLN_TRY((gemm<true, false, rlnp::kEpiF32>(p, dz, L.cout, 1, wbf(p, L.w_off), L.Kp, 1, M, L.Kp, L.cout,
                                                       nullptr, p.dtmp.as<float>(), nullptr, L.Kp)));
            const long long nx = (long long)n * L.H * L.H * L.cin;
            hipLaunchKernelGGL(rlnp::k_col2im_mask, dim3(grid1(nx)), dim3(rlnp::kB), 0, p.stream,
                               p.dtmp.as<const float>(), n, L.H, L.cin, L.stride, L.Ho, L.Kp,
                               p.x[l].as<const rlnp::bf16>(), dz_next);
            HIP_TRY(hipGetLastError());
        }
        std::swap(dz, dz_next);
    }
    if (p.red_pending) {  // the gradient is whole once the last reduction is done
        HIP_TRY(hipStreamWaitEvent(p.stream, p.ev_red, 0));
        p.red_pending = false;
    }
    return RLN_OK;
}

void col_sum(rln_proposal &p, const rlnp::bf16 *Z, int M, int N, float *out) {
    if (N % 8 == 0 && N <= 8 * rlnp::kB) {
        const int rows = std::max(256, (M + 2047) / 2048);
        hipLaunchKernelGGL(rlnp::k_col_sum8, dim3(1, (unsigned)((M + rows - 1) / rows)), dim3(rlnp::kB), 0, p.stream,
                           Z, M, N, rows, out);
    } else {
        const int rows = std::max(16, (M + 63) / 64);  // at most 64 row blocks
        hipLaunchKernelGGL(rlnp::k_col_sum, dim3(grid1(N), (M + rows - 1) / rows), dim3(rlnp::kB), 0, p.stream, Z,
                           M, N, rows, out);
    }
}

int adam_step(rln_proposal &p, double lr) {
    ++p.adam_t;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1. / (1. - std::pow(0.9, p.adam_t))), c2 = (float)(1. / (1. - std::pow(0.999, p.adam_t)));
    hipLaunchKernelGGL(rlnp::k_adam, dim3(grid1((long long)p.n_params)), dim3(rlnp::kB), 0, p.stream, p.w.as<float>(),
                       p.g.as<float>(), p.m1.as<float>(), p.m2.as<float>(), p.wb.as<rlnp::bf16>(),
                       (long long)p.n_params, (float)lr, b1, b2, eps, c1, c2);
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

int refresh_bf16(rln_proposal &p) {
    hipLaunchKernelGGL(rlnp::k_to_bf16, dim3(grid1((long long)p.n_params)), dim3(rlnp::kB), 0, p.stream,
                       p.w.as<const float>(), (long long)p.n_params, p.wb.as<rlnp::bf16>());
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

// network input images of n particle images (half-complex at current_size) into dst (bf16), optional FP32
int particle_images(rln_proposal &p, int n, const float *fimg, const float *fctf, int current_size, int on_device,
                    rlnp::bf16 *dst, float *dst_f32) {
    if (current_size < p.W || (current_size & 1)) RLN_FAIL(RLN_ERR_ARG, "current_size %d below the window %d", current_size, p.W);
    const int xs = current_size / 2 + 1;
    const size_t per = (size_t)current_size * xs;
    const float2 *srcp;
    const float *ctfp = nullptr;
    if (on_device) {
        srcp = reinterpret_cast<const float2 *>(fimg);
        ctfp = fctf;
    } else {
        RLN_TRY(p.src.ensure(sizeof(float2) * per * n));
        HIP_TRY(hipMemcpyAsync(p.src.ptr, fimg, sizeof(float2) * per * n, hipMemcpyHostToDevice, p.stream));
        srcp = p.src.as<const float2>();
        if (fctf) {
            RLN_TRY(p.src_ctf.ensure(sizeof(float) * per * n));
            HIP_TRY(hipMemcpyAsync(p.src_ctf.ptr, fctf, sizeof(float) * per * n, hipMemcpyHostToDevice, p.stream));
            ctfp = p.src_ctf.as<const float>();
        }
    }
    hipLaunchKernelGGL(rlnp::k_window_image, dim3(n), dim3(rlnp::kB), 0, p.stream, srcp, ctfp, (long long)per,
                       current_size, xs, p.W, dst, dst_f32);
    HIP_TRY(hipGetLastError());
    return RLN_OK;
}

int grow_training(rln_proposal &p, long long add) {
    const long long need = p.n_train + add;
    if (need <= p.train_cap) return RLN_OK;
    const long long cap = std::max(need, p.train_cap * 2);
    rln::DevBuf nb;
    RLN_TRY(nb.ensure(sizeof(rlnp::bf16) * (size_t)cap * p.W * p.W));
    if (p.n_train)
        HIP_TRY(hipMemcpyAsync(nb.ptr, p.train_img.ptr, sizeof(rlnp::bf16) * (size_t)p.n_train * p.W * p.W,
                               hipMemcpyDeviceToDevice, p.stream));
    HIP_TRY(hipStreamSynchronize(p.stream));
    std::swap(p.train_img.ptr, nb.ptr);
    std::swap(p.train_img.bytes, nb.bytes);
    p.train_cap = cap;
    return RLN_OK;
}

int pose_bin(const rln_proposal &p, double rot, double tilt, double psi) {
    const double a = rot * kPi / 180., b = tilt * kPi / 180.;
    const double z = std::cos(b);
    const double phi = std::atan2(std::sin(b) * std::sin(a), std::sin(b) * std::cos(a));
    const long long d = hp_ang2pix_nest(p.cfg.healpix_order, z, phi);
    const double step = 360. / p.cfg.n_psi_bins;
    int ip = (int)std::floor(psi / step + 0.5);
    ip %= p.cfg.n_psi_bins;
    if (ip < 0) ip += p.cfg.n_psi_bins;
    return (int)d * p.cfg.n_psi_bins + ip;
}

}  // namespace
}  // namespace rln

using namespace rln;

extern "C" int rln_proposal_create(int device, const rln_proposal_config *cfg, rln_proposal **out) {
    if (!cfg || !out) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_create: null argument");
    *out = nullptr;
    const int W = cfg->window, C = cfg->channels;
    if (W < 16 || W > rlnp::kMaxW || W % 8 || C < 8 || C % 8 || cfg->healpix_order < 0 || cfg->healpix_order > 6 ||
        cfg->n_psi_bins < 1)
        RLN_FAIL(RLN_ERR_ARG, "rln_proposal_create: unsupported configuration (window %d, channels %d)", W, C);
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) RLN_FAIL(RLN_ERR_ARG, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    rln_proposal *p = new rln_proposal();
    p->device = device;
    if (hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) p->n_cu = 0;
    p->cfg = *cfg;
    p->W = W;
    p->C = C;
    p->n_dir_bins = 12 << (2 * cfg->healpix_order);
    p->NB = p->n_dir_bins * cfg->n_psi_bins;
    p->F = C * (W / 8) * (W / 8);
    size_t off = 0;
    int H = W, cin = 1;
    for (int l = 0; l < kLayers; ++l) {
        ConvLayer &L = p->conv[l];
        L.cin = cin;
        L.cout = C;
        L.H = H;
        L.stride = kStride[l];
        L.Ho = (H - 1) / L.stride + 1;
        L.Kp = (cin == 1) ? 16 : 9 * cin;
        L.w_off = off;
        off += (size_t)C * L.Kp;
        L.b_off = off;
        off += C;
        H = L.Ho;
        cin = C;
    }
    p->fcw_off = off;
    off += (size_t)p->NB * p->F;
    p->fcb_off = off;
    off += p->NB;
    p->n_params = off;
    p->rng.seed(cfg->seed);
    int rc = RLN_OK;
    auto fail = [&](int r) {
        if (p->stream) (void)hipStreamDestroy(p->stream);
        if (p->side) (void)hipStreamDestroy(p->side);
        if (p->ev_dw) (void)hipEventDestroy(p->ev_dw);
        if (p->ev_red) (void)hipEventDestroy(p->ev_red);
        delete p;
        return r;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_dw, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_red, hipEventDisableTiming) != hipSuccess)
        return fail((set_error("hipStreamCreate / hipEventCreate failed"), RLN_ERR_HIP));
    const size_t nb = sizeof(float) * p->n_params;
    if ((rc = p->w.ensure(nb)) || (rc = p->g.ensure(nb)) || (rc = p->m1.ensure(nb)) || (rc = p->m2.ensure(nb)) ||
        (rc = p->wb.ensure(sizeof(rlnp::bf16) * p->n_params)))
        return fail(rc);
    // He-normal convolutions, zero biases, N(0, 1/F) fully connected layer
    std::vector<float> w(p->n_params, 0.f);
    std::normal_distribution<double> nd(0., 1.);
    for (int l = 0; l < kLayers; ++l) {
        const ConvLayer &L = p->conv[l];
        const double sd = std::sqrt(2. / (9. * L.cin));
        for (int o = 0; o < L.cout; ++o)
            for (int k = 0; k < 9 * L.cin; ++k) w[L.w_off + (size_t)o * L.Kp + k] = (float)(sd * nd(p->rng));
    }
    const double sfc = std::sqrt(1. / p->F);
    for (size_t i = 0; i < (size_t)p->NB * p->F; ++i) w[p->fcw_off + i] = (float)(sfc * nd(p->rng));
    if (hipMemcpy(p->w.ptr, w.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(p->g.ptr, 0, nb) != hipSuccess || hipMemset(p->m1.ptr, 0, nb) != hipSuccess ||
        hipMemset(p->m2.ptr, 0, nb) != hipSuccess)
        return fail((set_error("initialising the proposal weights failed"), RLN_ERR_HIP));
    if ((rc = refresh_bf16(*p)) != RLN_OK) return fail(rc);
    if (hipStreamSynchronize(p->stream) != hipSuccess) return fail((set_error("hipStreamSynchronize failed"), RLN_ERR_HIP));
    *out = p;
    return RLN_OK;
}

extern "C" int rln_proposal_destroy(rln_proposal *p) {
    if (!p) return RLN_OK;
    (void)hipSetDevice(p->device);
    if (p->stream) {
        (void)hipStreamSynchronize(p->stream);
        (void)hipStreamDestroy(p->stream);
    }
    if (p->side) {
        (void)hipStreamSynchronize(p->side);
        (void)hipStreamDestroy(p->side);
    }
    if (p->ev_dw) (void)hipEventDestroy(p->ev_dw);
    if (p->ev_red) (void)hipEventDestroy(p->ev_red);
    if (p->pend_ev) (void)hipEventDestroy(p->pend_ev);
    if (p->h_top) (void)hipHostFree(p->h_top);
    delete p;
    return RLN_OK;
}

extern "C" int rln_proposal_num_params(const rln_proposal *p, long long *n) {
    if (!p || !n) RLN_FAIL(RLN_ERR_ARG, "null argument");
    *n = (long long)p->n_params;
    return RLN_OK;
}

extern "C" int rln_proposal_num_bins(const rln_proposal *p, int *n) {
    if (!p || !n) RLN_FAIL(RLN_ERR_ARG, "null argument");
    *n = p->NB;
    return RLN_OK;
}

extern "C" int rln_proposal_get_params(rln_proposal *p, float *out) {
    if (!p || !out) RLN_FAIL(RLN_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(out, p->w.ptr, sizeof(float) * p->n_params, hipMemcpyDeviceToHost));
    return RLN_OK;
}

extern "C" int rln_proposal_set_params(rln_proposal *p, const float *in) {
    if (!p || !in) RLN_FAIL(RLN_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(p->w.ptr, in, sizeof(float) * p->n_params, hipMemcpyHostToDevice));
    RLN_TRY(refresh_bf16(*p));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return RLN_OK;
}

extern "C" int rln_proposal_clear_training(rln_proposal *p) {
    if (!p) RLN_FAIL(RLN_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    p->n_train = 0;
    p->train_lab.clear();
    return RLN_OK;
}

extern "C" int rln_proposal_training_size(const rln_proposal *p, long long *n) {
    if (!p || !n) RLN_FAIL(RLN_ERR_ARG, "null argument");
    *n = p->n_train;
    return RLN_OK;
}

extern "C" int rln_proposal_generate(rln_proposal *p, const rln_projector *ref, const rln_proposal_gen *gen, int n,
                                     long long first, double *poses) {
    if (!p || !ref || !gen || n < 0 || !gen->sigma2_noise) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_generate: null argument");
    if (!ref->data.ptr) RLN_FAIL(RLN_ERR_STATE, "rln_proposal_generate: reference not initialised");
    if (ref->device != p->device) RLN_FAIL(RLN_ERR_ARG, "reference on another device");
    if (gen->ori_size < p->W || !(gen->pixel_size > 0.)) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_generate: invalid box");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    const MdlDims &d = ref->dims;
    const int W = p->W, xw = W / 2 + 1;
    // the window must lie inside the reference's projection radius
    const int proj_r2 = d.maxr * d.maxr;
    if (W / 2 > d.maxr) RLN_FAIL(RLN_ERR_ARG, "window %d exceeds the reference radius %d", W, d.maxr);
    const int nshell = gen->ori_size / 2 + 1;
    std::vector<float> s2(nshell);
    for (int i = 0; i < nshell; ++i) s2[i] = (float)gen->sigma2_noise[i];
    DevBuf d_s2, d_smp, d_win;
    RLN_TRY(d_s2.ensure(sizeof(float) * nshell));
    HIP_TRY(hipMemcpyAsync(d_s2.ptr, s2.data(), sizeof(float) * nshell, hipMemcpyHostToDevice, p->stream));
    // CTF constants (src/ctf.cpp:134-261 with Q0 and no B-factor)
    const double kv = gen->voltage * 1e3;
    const double lambda = 12.2643247 / std::sqrt(kv * (1. + kv * 0.978466e-6));
    rlnp::CtfConst ck;
    ck.K1 = (float)(kPi * lambda);
    ck.K2 = (float)(kPi / 2. * gen->cs * 1e7 * lambda * lambda * lambda);
    ck.K3 = (float)std::atan(gen->q0 / std::sqrt(1. - gen->q0 * gen->q0));
    ck.inv_box_A = (float)(1. / (gen->ori_size * gen->pixel_size));
    ck.ori_size = (float)gen->ori_size;
    const rlnk::VolGeom vg{d.xdim, d.ydim, d.zdim, d.inity, d.initz};
    RLN_TRY(grow_training(*p, n));
    const int chunk = 4096;
    RLN_TRY(d_smp.ensure(sizeof(rlnp::GenSample) * chunk));
    RLN_TRY(d_win.ensure(sizeof(float2) * (size_t)chunk * W * xw));
    std::vector<rlnp::GenSample> smp(chunk);
    std::uniform_real_distribution<double> U(0., 1.);
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        HIP_TRY(hipStreamSynchronize(p->stream));  // smp is reused
        for (int i = 0; i < m; ++i) {
            // uniform rotation from a uniform unit quaternion (Shoemake 1992), then ZYZ angles
            const double u1 = U(p->rng), u2 = U(p->rng), u3 = U(p->rng);
            const double qw = std::sqrt(1. - u1) * std::sin(2. * kPi * u2), qx = std::sqrt(1. - u1) * std::cos(2. * kPi * u2);
            const double qy = std::sqrt(u1) * std::sin(2. * kPi * u3), qz = std::sqrt(u1) * std::cos(2. * kPi * u3);
            const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                                 2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                                 2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
            double rot, tilt, psi;
            matrix_to_angles(R, rot, tilt, psi);
            double A[9];
            euler_angles2matrix(rot, tilt, psi, A);
            rlnp::GenSample &s = smp[i];
            s.fr = rlnk::Frame{(float)(d.pad * A[0]), (float)(d.pad * A[1]), (float)(d.pad * A[2]),
                               (float)(d.pad * A[3]), (float)(d.pad * A[4]), (float)(d.pad * A[5])};
            s.dx = (float)((2. * U(p->rng) - 1.) * gen->offset_max);
            s.dy = (float)((2. * U(p->rng) - 1.) * gen->offset_max);
            const double du = gen->defocus_min + U(p->rng) * (gen->defocus_max - gen->defocus_min);
            s.defU = (float)du;
            s.defV = (float)(du - U(p->rng) * gen->max_astigmatism);
            s.defAng = (float)(180. * U(p->rng));
            const double ps = psi - 360. * std::floor(psi / 360.);
            p->train_lab.push_back(pose_bin(*p, rot, tilt, ps));
            if (poses) {
                poses[3 * (size_t)(i0 + i)] = rot;
                poses[3 * (size_t)(i0 + i) + 1] = tilt;
                poses[3 * (size_t)(i0 + i) + 2] = ps;
            }
        }
        HIP_TRY(hipMemcpyAsync(d_smp.ptr, smp.data(), sizeof(rlnp::GenSample) * m, hipMemcpyHostToDevice, p->stream));
        hipLaunchKernelGGL(rlnp::k_gen_window, dim3(grid1((long long)W * xw), m), dim3(rlnp::kB), 0, p->stream,
                           ref->data.as<const float4>(), vg, proj_r2, d_smp.as<const rlnp::GenSample>(), ck,
                           d_s2.as<const float>(), nshell, W, (unsigned long long)p->cfg.seed, first + i0,
                           d_win.as<float2>());
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(rlnp::k_window_image, dim3(m), dim3(rlnp::kB), 0, p->stream, d_win.as<const float2>(),
                           (const float *)nullptr, (long long)W * xw, W, xw, W,
                           p->train_img.as<rlnp::bf16>() + (size_t)(p->n_train + i0) * W * W, (float *)nullptr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->n_train += n;
    return RLN_OK;
}

extern "C" int rln_proposal_add_images(rln_proposal *p, int n, const float *fimg, const float *fctf,
                                       int current_size, int on_device, const double *poses) {
    if (!p || n < 0 || (n > 0 && (!fimg || !poses))) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_add_images: null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    RLN_TRY(grow_training(*p, n));
    const int chunk = 2048;
    const size_t per = (size_t)current_size * (current_size / 2 + 1);
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        RLN_TRY(particle_images(*p, m, fimg + 2 * per * i0, fctf ? fctf + per * i0 : nullptr, current_size, on_device,
                                p->train_img.as<rlnp::bf16>() + (size_t)(p->n_train + i0) * p->W * p->W, nullptr));
        HIP_TRY(hipStreamSynchronize(p->stream));
    }
    for (int i = 0; i < n; ++i) {
        const double ps = poses[3 * i + 2] - 360. * std::floor(poses[3 * i + 2] / 360.);
        p->train_lab.push_back(pose_bin(*p, poses[3 * i], poses[3 * i + 1], ps));
    }
    p->n_train += n;
    return RLN_OK;
}

extern "C" int rln_proposal_train(rln_proposal *p, int epochs, int batch, double lr, double *mean_loss) {
    if (!p || epochs < 0 || batch <= 0 || !(lr > 0.)) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_train: invalid argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    if (p->n_train < batch) RLN_FAIL(RLN_ERR_STATE, "training set (%lld) smaller than one batch (%d)", p->n_train, batch);
    HIP_TRY(hipSetDevice(p->device));
    RLN_TRY(ensure_batch(*p, batch));
    const long long N = p->n_train;
    const long long steps = N / batch;
    DevBuf d_lab, d_losses;
    RLN_TRY(d_lab.ensure(sizeof(int) * N));
    RLN_TRY(d_losses.ensure(sizeof(double) * steps * batch));
    HIP_TRY(hipMemcpyAsync(d_lab.ptr, p->train_lab.data(), sizeof(int) * N, hipMemcpyHostToDevice, p->stream));
    std::vector<int> order(N);
    std::vector<int> lab_batch(batch);
    DevBuf d_order;
    RLN_TRY(d_order.ensure(sizeof(int) * N));
    double last = 0.;
    for (int ep = 0; ep < epochs; ++ep) {
        for (long long i = 0; i < N; ++i) order[i] = (int)i;
        std::shuffle(order.begin(), order.end(), p->rng);
        HIP_TRY(hipStreamSynchronize(p->stream));
        HIP_TRY(hipMemcpyAsync(d_order.ptr, order.data(), sizeof(int) * N, hipMemcpyHostToDevice, p->stream));
        for (long long s = 0; s < steps; ++s) {
            const int *ib = d_order.as<const int>() + s * batch;
            hipLaunchKernelGGL(rlnp::k_gather_images, dim3(grid1((long long)p->W * p->W), batch), dim3(rlnp::kB), 0,
                               p->stream, p->train_img.as<const rlnp::bf16>(), ib, p->W * p->W, p->x[0].as<rlnp::bf16>());
            HIP_TRY(hipGetLastError());
            // labels of the batch: gathered on the device from the label array
            hipLaunchKernelGGL(rlnp::k_gather_labels, dim3(grid1(batch)), dim3(rlnp::kB), 0, p->stream,
                               d_lab.as<const int>(), ib, batch, p->labels.as<int>());
            HIP_TRY(hipGetLastError());
            RLN_TRY(forward(*p, batch));
            RLN_TRY(backward(*p, batch));
            HIP_TRY(hipMemcpyAsync(d_losses.as<double>() + s * batch, p->loss.ptr, sizeof(double) * batch,
                                   hipMemcpyDeviceToDevice, p->stream));
            RLN_TRY(adam_step(*p, lr));
        }
        std::vector<double> ls((size_t)steps * batch);
        HIP_TRY(hipMemcpyAsync(ls.data(), d_losses.ptr, sizeof(double) * ls.size(), hipMemcpyDeviceToHost, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        double sum = 0.;
        for (double v : ls) sum += v;
        last = ls.empty() ? 0. : sum / (double)ls.size();
    }
    if (mean_loss) *mean_loss = last;
    return RLN_OK;
}

extern "C" int rln_proposal_predict_launch(rln_proposal *p, int n, const float *fimg, const float *fctf,
                                           int current_size, int on_device, int topk) {
    if (!p || n < 0 || (n > 0 && !fimg)) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_predict_launch: null argument");
    if (topk < 1 || topk > rlnp::kMaxTopK || topk > p->NB) RLN_FAIL(RLN_ERR_ARG, "topk %d out of range", topk);
    std::lock_guard<std::mutex> lk(p->mtx);
    if (p->pend_n >= 0) RLN_FAIL(RLN_ERR_STATE, "a prediction is in flight: finish it first");
    HIP_TRY(hipSetDevice(p->device));
    if (!p->pend_ev) HIP_TRY(hipEventCreateWithFlags(&p->pend_ev, hipEventDisableTiming));
    const size_t need = (sizeof(int) + sizeof(float)) * (size_t)std::max(n, 1) * topk;
    if (need > p->h_top_bytes) {
        if (p->h_top) HIP_TRY(hipHostFree(p->h_top));
        p->h_top = nullptr;
        p->h_top_bytes = 0;
        HIP_TRY(hipHostMalloc(&p->h_top, need, hipHostMallocDefault));
        p->h_top_bytes = need;
    }
    // chunks run back to back on the stream (they share the activation buffers in stream order); the top-k
    // of all particles collect on the device and come back in one copy into pinned memory
    const int chunk = 2048;
    RLN_TRY(ensure_batch(*p, std::min(n, chunk)));
    RLN_TRY(p->top_bins.ensure(sizeof(int) * (size_t)std::max(n, 1) * topk));
    RLN_TRY(p->top_probs.ensure(sizeof(float) * (size_t)std::max(n, 1) * topk));
    const size_t per = (size_t)current_size * (current_size / 2 + 1);
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        RLN_TRY(particle_images(*p, m, fimg + 2 * per * i0, fctf ? fctf + per * i0 : nullptr, current_size, on_device,
                                p->x[0].as<rlnp::bf16>(), nullptr));
        RLN_TRY(forward(*p, m, false));
        hipLaunchKernelGGL(rlnp::k_topk, dim3(m), dim3(rlnp::kB), sizeof(float) * p->NB, p->stream,
                           p->logits.as<const float>(), p->NB, topk, p->top_bins.as<int>() + (size_t)i0 * topk,
                           p->top_probs.as<float>() + (size_t)i0 * topk);
        HIP_TRY(hipGetLastError());
    }
    if (n > 0) {
        char *h = static_cast<char *>(p->h_top);
        HIP_TRY(hipMemcpyAsync(h, p->top_bins.ptr, sizeof(int) * (size_t)n * topk, hipMemcpyDeviceToHost, p->stream));
        HIP_TRY(hipMemcpyAsync(h + sizeof(int) * (size_t)n * topk, p->top_probs.ptr, sizeof(float) * (size_t)n * topk,
                               hipMemcpyDeviceToHost, p->stream));
    }
    HIP_TRY(hipEventRecord(p->pend_ev, p->stream));
    p->pend_n = n;
    p->pend_topk = topk;
    return RLN_OK;
}

extern "C" int rln_proposal_predict_finish(rln_proposal *p, int *bins, float *probs) {
    if (!p) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_predict_finish: null proposal");
    std::lock_guard<std::mutex> lk(p->mtx);
    if (p->pend_n < 0) RLN_FAIL(RLN_ERR_STATE, "no prediction in flight");
    const size_t k = (size_t)p->pend_n * p->pend_topk;
    if (k > 0 && (!bins || !probs)) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_predict_finish: null output");
    HIP_TRY(hipSetDevice(p->device));
    p->pend_n = -1;  // the in-flight state ends here whatever the outcome
    HIP_TRY(hipEventSynchronize(p->pend_ev));
    if (k > 0) {
        const char *h = static_cast<const char *>(p->h_top);
        std::memcpy(bins, h, sizeof(int) * k);
        std::memcpy(probs, h + sizeof(int) * k, sizeof(float) * k);
    }
    return RLN_OK;
}

extern "C" int rln_proposal_predict(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size,
                                    int on_device, int topk, int *bins, float *probs) {
    if (!p || n < 0 || (n > 0 && (!fimg || !bins || !probs))) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_predict: null argument");
    RLN_TRY(rln_proposal_predict_launch(p, n, fimg, fctf, current_size, on_device, topk));
    return rln_proposal_predict_finish(p, bins, probs);
}

extern "C" int rln_proposal_pose_bin(const rln_proposal *p, double rot, double tilt, double psi, int *bin) {
    if (!p || !bin) RLN_FAIL(RLN_ERR_ARG, "null argument");
    *bin = pose_bin(*p, rot, tilt, psi - 360. * std::floor(psi / 360.));
    return RLN_OK;
}

namespace {
// grid orientations of one bin: NEST children of its direction x grid psi within half a bin of its centre
void bin_orients(const rln_proposal &p, int grid_order, int n_psi_grid, int bin, std::vector<int> &out) {
    out.clear();
    const int d = bin / p.cfg.n_psi_bins, ib = bin % p.cfg.n_psi_bins;
    const int nchild = 1 << (2 * (grid_order - p.cfg.healpix_order));
    const double sb = 360. / p.cfg.n_psi_bins, sg = 360. / n_psi_grid;
    std::vector<int> psis;
    for (int j = 0; j < n_psi_grid; ++j) {
        double dd = std::fmod(j * sg - ib * sb, 360.);
        if (dd > 180.) dd -= 360.;
        if (dd < -180.) dd += 360.;
        if (std::fabs(dd) <= 0.5 * sb + 1e-9) psis.push_back(j);
    }
    for (int k = 0; k < nchild; ++k)
        for (int j : psis) out.push_back((d * nchild + k) * n_psi_grid + j);
}
}  // namespace

extern "C" int rln_proposal_bin_orientations(const rln_proposal *p, int grid_order, int n_psi_grid, int bin, int *out,
                                             int capacity, int *count) {
    if (!p || !count || (capacity > 0 && !out)) RLN_FAIL(RLN_ERR_ARG, "null argument");
    if (grid_order < p->cfg.healpix_order || grid_order > 12 || n_psi_grid < 1 || bin < 0 || bin >= p->NB)
        RLN_FAIL(RLN_ERR_ARG, "rln_proposal_bin_orientations: invalid argument");
    std::vector<int> o;
    bin_orients(*p, grid_order, n_psi_grid, bin, o);
    for (int c = 0; c < (int)o.size() && c < capacity; ++c) out[c] = o[c];
    *count = (int)o.size();
    return RLN_OK;
}

extern "C" int rln_proposal_orientation_sets(const rln_proposal *p, int n, const int *bins, const float *probs,
                                             int topk, int grid_order, int n_psi_grid, double min_top1,
                                             int *set_offsets, int *set_orients, long long capacity) {
    if (!p || n < 0 || (n > 0 && (!bins || !probs || !set_offsets || !set_orients)) || topk < 1)
        RLN_FAIL(RLN_ERR_ARG, "rln_proposal_orientation_sets: null argument");
    if (grid_order < p->cfg.healpix_order || grid_order > 12 || n_psi_grid < 1)
        RLN_FAIL(RLN_ERR_ARG, "rln_proposal_orientation_sets: invalid grid");
    for (long long q = 0; q < (long long)n * topk; ++q)
        if (bins[q] < 0 || bins[q] >= p->NB) RLN_FAIL(RLN_ERR_ARG, "bin %d out of range", bins[q]);
    // orientations of every bin (built once per grid), then the particles' sorted unions in parallel; offsets and
    // the copy into the caller's arrays follow in particle order
    std::lock_guard<std::mutex> lk(p->orient_mtx);
    if (p->orient_order != grid_order || p->orient_npsi != n_psi_grid || (int)p->orient_cache.size() != p->NB) {
        p->orient_cache.assign(p->NB, std::vector<int>());
        for (int b = 0; b < p->NB; ++b) bin_orients(*p, grid_order, n_psi_grid, b, p->orient_cache[b]);
        p->orient_order = grid_order;
        p->orient_npsi = n_psi_grid;
    }
    const std::vector<std::vector<int>> &cache = p->orient_cache;
    std::vector<std::vector<int>> lists(n);
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto build = [&](int t0) {
        std::vector<int> acc;
        for (int i = t0; i < n; i += nt) {
            acc.clear();
            if (probs[(size_t)i * topk] >= min_top1) {
                for (int t = 0; t < topk; ++t) {
                    const std::vector<int> &c = cache[bins[(size_t)i * topk + t]];
                    acc.insert(acc.end(), c.begin(), c.end());
                }
                std::sort(acc.begin(), acc.end());
                acc.erase(std::unique(acc.begin(), acc.end()), acc.end());
            }
            lists[i] = acc;
        }
    };
    {
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(build, t);
        build(0);
        for (auto &th : pool) th.join();
    }
    long long total = 0;
    set_offsets[0] = 0;
    for (int i = 0; i < n; ++i) {
        total += (long long)lists[i].size();
        if (total > capacity) RLN_FAIL(RLN_ERR_ARG, "orientation list capacity %lld exceeded", capacity);
        if (total > 0x7fffffffLL) RLN_FAIL(RLN_ERR_ARG, "orientation lists exceed 2^31 entries");
        set_offsets[i + 1] = (int)total;
    }
    for (int i = 0; i < n; ++i) std::copy(lists[i].begin(), lists[i].end(), set_orients + set_offsets[i]);
    return RLN_OK;
}

extern "C" int rln_proposal_images(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size,
                                   int on_device, float *out) {
    if (!p || n < 0 || (n > 0 && (!fimg || !out))) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_images: null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    RLN_TRY(ensure_batch(*p, n));
    DevBuf f;
    RLN_TRY(f.ensure(sizeof(float) * (size_t)n * p->W * p->W));
    RLN_TRY(particle_images(*p, n, fimg, fctf, current_size, on_device, p->x[0].as<rlnp::bf16>(), f.as<float>()));
    HIP_TRY(hipMemcpyAsync(out, f.ptr, sizeof(float) * (size_t)n * p->W * p->W, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return RLN_OK;
}

extern "C" int rln_proposal_forward_backward(rln_proposal *p, int n, const float *images, const int *labels,
                                             double *loss, float *grads, float *logits) {
    if (!p || n <= 0 || !images || !labels) RLN_FAIL(RLN_ERR_ARG, "rln_proposal_forward_backward: null argument");
    std::lock_guard<std::mutex> lk(p->mtx);
    HIP_TRY(hipSetDevice(p->device));
    for (int i = 0; i < n; ++i)
        if (labels[i] < 0 || labels[i] >= p->NB) RLN_FAIL(RLN_ERR_ARG, "label %d out of range", labels[i]);
    RLN_TRY(ensure_batch(*p, n));
    const size_t ni = (size_t)n * p->W * p->W;
    DevBuf f;
    RLN_TRY(f.ensure(sizeof(float) * ni));
    HIP_TRY(hipMemcpyAsync(f.ptr, images, sizeof(float) * ni, hipMemcpyHostToDevice, p->stream));
    hipLaunchKernelGGL(rlnp::k_to_bf16, dim3(grid1((long long)ni)), dim3(rlnp::kB), 0, p->stream, f.as<const float>(),
                       (long long)ni, p->x[0].as<rlnp::bf16>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(p->labels.ptr, labels, sizeof(int) * n, hipMemcpyHostToDevice, p->stream));
    // gradients start from zero here; the optimiser state is untouched
    HIP_TRY(hipMemsetAsync(p->g.ptr, 0, sizeof(float) * p->n_params, p->stream));
    RLN_TRY(forward(*p, n));
    RLN_TRY(backward(*p, n));
    std::vector<double> ls(n);
    HIP_TRY(hipMemcpyAsync(ls.data(), p->loss.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
    if (grads) HIP_TRY(hipMemcpyAsync(grads, p->g.ptr, sizeof(float) * p->n_params, hipMemcpyDeviceToHost, p->stream));
    if (logits)
        HIP_TRY(hipMemcpyAsync(logits, p->logits.ptr, sizeof(float) * (size_t)n * p->NB, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipMemsetAsync(p->g.ptr, 0, sizeof(float) * p->n_params, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (loss) {
        double s = 0.;
        for (double v : ls) s += v;
        *loss = s / n;
    }
    return RLN_OK;
}
