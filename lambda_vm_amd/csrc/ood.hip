// Out-of-domain stage kernels for gfx950: Fp3 batch inversion, inverted
// denominators over the LDE coset, barycentric evaluation of resident LDE
// columns at K points, and the DEEP composition codeword.
//
// Coset points: physical row p of a resident LDE table (rows stored in
// bit-reversed order) is the natural LDE index bitrev_N(p), i.e. the point
// x_p = g * w_N^bitrev_N(p). Its first n physical rows are the sub-coset
// g*<w_n> the barycentric formula of crypto/stark/src/trace.rs:831-1110
// reads (natural rows i*B), so those reads are contiguous here.
#include <cstring>
#include <vector>

#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

__device__ __forceinline__ fp3 ld3(const uint64_t *p) { return fp3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void st3(uint64_t *p, fp3 v) {
    p[0] = v.c0;
    p[1] = v.c1;
    p[2] = v.c2;
}

// Montgomery batch inversion (crypto/math/src/field/element.rs:64-111) with
// interleaved ownership: thread t inverts elements t, t+T, t+2T, ... so each
// wave's loads and stores stay coalesced; one Fp3 inversion per thread.
// `out` holds the running prefix products during the forward sweep.
__global__ __launch_bounds__(256) void batch_inverse_fp3_kernel(const uint64_t *__restrict__ in,
                                                                uint64_t *__restrict__ out, uint64_t n) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    fp3 acc = fp3{1, 0, 0};
    uint64_t i = t;
    for (; i < n; i += T) {
        acc = f3_mul(acc, ld3(in + 3 * i));
        st3(out + 3 * i, acc);
    }
    fp3 inv_acc = f3_inv(acc);
    // walk back: out[i] = inv_acc * prefix[i - T]; inv_acc *= in[i]
    uint64_t last = t + ((n - 1 - t) / T) * T;
    for (i = last;; i -= T) {
        if (i >= T) {
            fp3 prev = ld3(out + 3 * (i - T));
            fp3 xi = ld3(in + 3 * i);
            st3(out + 3 * i, f3_mul(inv_acc, prev));
            inv_acc = f3_mul(inv_acc, xi);
        } else {
            st3(out + 3 * i, inv_acc);
            break;
        }
    }
}

// den[k * rows + p] = sign ? (z_k - x_p) : (x_p - z_k) for the first `rows`
// physical rows of an LDE of 2^log_N rows on offset*<w_N>.
__global__ void coset_denoms_kernel(uint64_t *den, uint64_t rows, uint32_t log_N, uint64_t offset,
                                    const uint64_t *z, uint32_t K, int z_minus_x, const uint64_t *lo,
                                    const uint64_t *hi) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= rows) return;
    uint32_t e = log_N ? (bitrev((uint32_t)p, log_N) << (32 - log_N)) : 0u;
    uint64_t x = mul(offset, mul(lo[e & 0xFFFFu], hi[e >> 16]));
    for (uint32_t k = 0; k < K; k++) {
        fp3 zk = fp3{z[3 * k], z[3 * k + 1], z[3 * k + 2]};
        fp3 d = z_minus_x ? fp3{sub(zk.c0, x), zk.c1, zk.c2} : fp3{sub(x, zk.c0), neg(zk.c1), neg(zk.c2)};
        st3(den + 3 * (k * rows + p), d);
    }
}

// scale[k * n + p] = x_p * inv_den[k * n + p]
__global__ void bary_scale_kernel(uint64_t *s, uint64_t n, uint32_t K, uint32_t log_N, uint64_t offset,
                                  const uint64_t *lo, const uint64_t *hi) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t e = log_N ? (bitrev((uint32_t)p, log_N) << (32 - log_N)) : 0u;
    uint64_t x = mul(offset, mul(lo[e & 0xFFFFu], hi[e >> 16]));
    for (uint32_t k = 0; k < K; k++) st3(s + 3 * (k * n + p), f3_mul_base(ld3(s + 3 * (k * n + p)), x));
}

constexpr int BARY_MAX_K = 8;
constexpr int BARY_ROWS_PER_BLOCK = 1024;

// partial[chunk][k][c] = sum over the chunk's rows p of lde[p][c] * scale[k][p]
// block: 256 threads = 4 row lanes x 64 columns; grid (chunks, column groups).
// The three components of each point's sum are delayed-reduction dot
// products (one 128-bit accumulate per product, one reduction per thread).
template <int K>
__global__ __launch_bounds__(256) void bary_partial_kernel(const uint64_t *__restrict__ lde, uint64_t pitch,
                                                           uint64_t w, uint64_t n, const uint64_t *__restrict__ scale,
                                                           uint64_t *__restrict__ partial) {
    __shared__ uint64_t red[4][K][64][3];
    const int lane = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const uint64_t c = (uint64_t)blockIdx.y * 64 + lane;
    const uint64_t p0 = (uint64_t)blockIdx.x * BARY_ROWS_PER_BLOCK;
    acc128 acc[K][3];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) acc_zero(acc[k][j]);
    if (c < w) {
        const uint64_t pend = p0 + BARY_ROWS_PER_BLOCK < n ? p0 + BARY_ROWS_PER_BLOCK : n;
        for (uint64_t p = p0 + rl; p < pend; p += 4) {
            uint64_t f = lde[p * pitch + c];
            // p is the same for the 64 lanes of a wave: read the scales through
            // the scalar path once per wave instead of 64 identical vector loads
            const uint64_t pu = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(p >> 32)) << 32) |
                                __builtin_amdgcn_readfirstlane((uint32_t)p);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const uint64_t *s = scale + 3 * (k * n + pu);
                acc_mad(acc[k][0], s[0], f);
                acc_mad(acc[k][1], s[1], f);
                acc_mad(acc[k][2], s[2], f);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) red[rl][k][lane][j] = acc_reduce(acc[k][j]);
    __syncthreads();
    if (rl == 0 && c < w) {
        for (int k = 0; k < K; k++) {
            fp3 s = fp3{red[0][k][lane][0], red[0][k][lane][1], red[0][k][lane][2]};
            for (int r = 1; r < 4; r++) s = f3_add(s, fp3{red[r][k][lane][0], red[r][k][lane][1], red[r][k][lane][2]});
            st3(partial + 3 * (((uint64_t)blockIdx.x * K + k) * w + c), s);
        }
    }
}

template <int K>
static void launch_bary_partial(dim3 grid, hipStream_t st, const uint64_t *lde, uint64_t w, uint64_t n,
                                const uint64_t *scale, uint64_t *partial) {
    hipLaunchKernelGGL(bary_partial_kernel<K>, grid, dim3(256), 0, st, lde, w, w, n, scale, partial);
}

// out[k][c] = vf[k] * sum over chunks of partial[chunk][k][c]
__global__ void bary_combine_kernel(const uint64_t *partial, uint64_t chunks, uint32_t K, uint64_t w,
                                    const uint64_t *vf, uint64_t *out) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)K * w) return;
    uint64_t k = idx / w, c = idx - k * w;
    fp3 s = fp3{0, 0, 0};
    for (uint64_t ch = 0; ch < chunks; ch++) s = f3_add(s, ld3(partial + 3 * ((ch * K + k) * w + c)));
    st3(out + 3 * idx, f3_mul(ld3(vf + 3 * k), s));
}

// DEEP composition at every LDE row (crypto/math-cuda/kernels/deep.cu:1-25,
// crypto/stark/src/prover.rs:2400-2604), in storage (bit-reversed) order:
//   deep[p] = inv_h[p] * (sum_j gh_j * H_j[p] - ch)
//           + sum_k inv_t[k][p] * (sum_j gt_{j,k} * t_j[p] - ct_k)
// where ch = sum_j gh_j * H_j(z^P) and ct_k = sum_j gt_{j,k} * t_j(z_k) are
// computed once on the host, main columns are base words, aux and parts are
// Fp3 (3 words each).
constexpr int DEEP_MAX_K = 4;
constexpr int DEEP_CHUNK = 16;  // main columns staged per LDS round
// block: 256 consecutive physical rows. The main (base) columns of the block
// are staged through LDS 16 at a time with coalesced row-segment loads; each
// thread accumulates sum_j gt_{j,k} * t_j[p] as 3K delayed-reduction dot
// products.
template <int K>
__global__ __launch_bounds__(256) void deep_kernel(uint64_t N, const uint64_t *__restrict__ main_lde, uint64_t wm,
                                                   const uint64_t *__restrict__ aux_lde, uint64_t wa,
                                                   const uint64_t *__restrict__ parts_lde, uint64_t np,
                                                   const uint64_t *__restrict__ gtr, const uint64_t *__restrict__ gh,
                                                   const uint64_t *__restrict__ consts,
                                                   const uint64_t *__restrict__ inv_den, uint64_t *__restrict__ out) {
    __shared__ uint64_t tile[256][DEEP_CHUNK + 1];
    __shared__ uint64_t sg[DEEP_CHUNK][K][3];
    const int t = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * 256;
    const uint64_t p = r0 + t;
    acc128 acc[K][3];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) acc_zero(acc[k][j]);
    for (uint64_t j0 = 0; j0 < wm; j0 += DEEP_CHUNK) {
        const int cw = (int)(wm - j0 < DEEP_CHUNK ? wm - j0 : DEEP_CHUNK);
        __syncthreads();
        for (int e = t; e < 256 * DEEP_CHUNK; e += 256) {
            const int rr = e / DEEP_CHUNK, cc = e % DEEP_CHUNK;
            tile[rr][cc] = (cc < cw && r0 + rr < N) ? main_lde[(r0 + rr) * wm + j0 + cc] : 0;
        }
        for (int e = t; e < DEEP_CHUNK * K * 3; e += 256) {
            const int cc = e / (K * 3), rem = e % (K * 3);
            sg[cc][rem / 3][rem % 3] = cc < cw ? gtr[3 * ((j0 + cc) * K + rem / 3) + rem % 3] : 0;
        }
        __syncthreads();
#pragma unroll 4
        for (int cc = 0; cc < DEEP_CHUNK; cc++) {
            const uint64_t v = tile[t][cc];
#pragma unroll
            for (int k = 0; k < K; k++) {
                acc_mad(acc[k][0], sg[cc][k][0], v);
                acc_mad(acc[k][1], sg[cc][k][1], v);
                acc_mad(acc[k][2], sg[cc][k][2], v);
            }
        }
    }
    if (p >= N) return;
    fp3 tk[K];
#pragma unroll
    for (int k = 0; k < K; k++) tk[k] = fp3{acc_reduce(acc[k][0]), acc_reduce(acc[k][1]), acc_reduce(acc[k][2])};
    if (aux_lde) {
        const uint64_t *arow = aux_lde + p * 3 * wa;
        for (uint64_t j = 0; j < wa; j++) {
            fp3 v = ld3(arow + 3 * j);
#pragma unroll
            for (int k = 0; k < K; k++) tk[k] = f3_add(tk[k], f3_mul(ld3(gtr + 3 * ((wm + j) * K + k)), v));
        }
    }
    fp3 r = fp3{0, 0, 0};
#pragma unroll
    for (int k = 0; k < K; k++)
        r = f3_add(r, f3_mul(ld3(inv_den + 3 * (k * N + p)), f3_sub(tk[k], ld3(consts + 3 * k))));
    if (parts_lde && np) {
        fp3 hs = fp3{0, 0, 0};
        const uint64_t *prow = parts_lde + p * 3 * np;
        for (uint64_t j = 0; j < np; j++) hs = f3_add(hs, f3_mul(ld3(gh + 3 * j), ld3(prow + 3 * j)));
        r = f3_add(r, f3_mul(ld3(inv_den + 3 * (K * N + p)), f3_sub(hs, ld3(consts + 3 * K))));
    }
    st3(out + 3 * p, r);
}

template <int K>
static void launch_deep(hipStream_t st, uint64_t N, const uint64_t *main_lde, uint64_t wm, const uint64_t *aux_lde,
                        uint64_t wa, const uint64_t *parts_lde, uint64_t np, const uint64_t *gtr, const uint64_t *gh,
                        const uint64_t *consts, const uint64_t *inv_den, uint64_t *out) {
    hipLaunchKernelGGL(deep_kernel<K>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, main_lde, wm, aux_lde,
                       wa, parts_lde, np, gtr, gh, consts, inv_den, out);
}

}  // namespace lvm

using namespace lvm;

namespace {
unsigned log2u(uint64_t x) { return (unsigned)__builtin_ctzll(x); }

struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch() {
        for (void *p : ptrs) lvm_free(p);
    }
    template <class T>
    int get(T **out, size_t bytes) {
        void *p = nullptr;
        hipError_t e = lvm_malloc(&p, bytes ? bytes : 8);
        if (e != hipSuccess) return fail(LVM_ERR_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
        ptrs.push_back(p);
        *out = (T *)p;
        return LVM_OK;
    }
};

#define TRYS(x)                        \
    do {                               \
        int _rc = (x);                 \
        if (_rc != LVM_OK) return _rc; \
    } while (0)

void launch_batch_inverse(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t st) {
    if (!n) return;
    // ~64 elements per thread amortise the one inversion each thread performs
    uint64_t threads = (n + 63) / 64;
    if (threads < 256) threads = n < 256 ? n : 256;
    uint64_t blocks = (threads + 255) / 256;
    hipLaunchKernelGGL(batch_inverse_fp3_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
}

// host Fp3 helpers
struct H3 {
    uint64_t c[3];
};
H3 h3mul(H3 a, H3 b) {
    fp3 r = f3_mul(fp3{a.c[0], a.c[1], a.c[2]}, fp3{b.c[0], b.c[1], b.c[2]});
    return H3{{r.c0, r.c1, r.c2}};
}
}  // namespace

extern "C" {

int lvm_batch_inverse_ext3_dev(const uint64_t *dev_in, uint64_t *dev_out, size_t n) {
    if ((!dev_in || !dev_out) && n) return fail(LVM_ERR_ARG, "null argument");
    if (dev_in == dev_out && n) return fail(LVM_ERR_ARG, "in and out must differ");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    {
        ProfScope ps("batch_inverse", thread_stream(d));
        launch_batch_inverse(dev_in, dev_out, n, thread_stream(d));
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_batch_inverse_ext3(const uint64_t *in, uint64_t *out, size_t n) {
    if ((!in || !out) && n) return fail(LVM_ERR_ARG, "null argument");
    if (!n) return LVM_OK;
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    Scratch s;
    uint64_t *di, *dout;
    TRYS(s.get(&di, n * 24));
    TRYS(s.get(&dout, n * 24));
    LVM_HIP(hipMemcpyAsync(di, in, n * 24, hipMemcpyHostToDevice, thread_stream(d)));
    int rc = lvm_batch_inverse_ext3_dev(di, dout, n);
    if (rc) return rc;
    LVM_HIP(hipMemcpyAsync(out, dout, n * 24, hipMemcpyDeviceToHost, thread_stream(d)));
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_barycentric_base_multi(const lvm_lde *h, uint64_t coset_offset, const uint64_t *z, size_t K, uint64_t *out) {
    if (!h || !z || !out || K == 0) return fail(LVM_ERR_ARG, "bad barycentric arguments");
    if (K > (size_t)BARY_MAX_K) return fail(LVM_ERR_ARG, "at most 8 evaluation points per call");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = thread_stream(d);
    const uint64_t n = h->n, w = h->w;
    const unsigned logN = log2u(h->N);
    Scratch s;
    uint64_t *dz, *den, *scale, *partial, *vfd, *dout;
    const uint64_t chunks = (n + BARY_ROWS_PER_BLOCK - 1) / BARY_ROWS_PER_BLOCK;
    TRYS(s.get(&dz, K * 24));
    TRYS(s.get(&den, K * n * 24));
    TRYS(s.get(&scale, K * n * 24));
    TRYS(s.get(&partial, chunks * K * w * 24));
    TRYS(s.get(&vfd, K * 24));
    TRYS(s.get(&dout, K * w * 24));
    // vanishing factors n^-1 * g^-n * (z_k^n - g^n) (trace.rs:900-905)
    std::vector<uint64_t> vf(3 * K);
    const uint64_t gn = pow(coset_offset, n), nig = mul(inv((uint64_t)n), inv(gn));
    for (size_t k = 0; k < K; k++) {
        fp3 zn = f3_pow(fp3{z[3 * k], z[3 * k + 1], z[3 * k + 2]}, n);
        fp3 v = f3_mul_base(fp3{sub(zn.c0, gn), zn.c1, zn.c2}, nig);
        vf[3 * k] = v.c0;
        vf[3 * k + 1] = v.c1;
        vf[3 * k + 2] = v.c2;
    }
    LVM_HIP(hipMemcpyAsync(dz, z, K * 24, hipMemcpyHostToDevice, st));
    LVM_HIP(hipMemcpyAsync(vfd, vf.data(), K * 24, hipMemcpyHostToDevice, st));
    {
        ProfScope ps("bary_denoms", st);
        hipLaunchKernelGGL(coset_denoms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, den, n, logN,
                           coset_offset, dz, (uint32_t)K, 1, d->tb.fwd_lo, d->tb.fwd_hi);
        launch_batch_inverse(den, scale, K * n, st);
        hipLaunchKernelGGL(bary_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, scale, n,
                           (uint32_t)K, logN, coset_offset, d->tb.fwd_lo, d->tb.fwd_hi);
    }
    {
        ProfScope ps("bary_sum", st);
        dim3 grid((unsigned)chunks, (unsigned)((w + 63) / 64));
        switch (K) {
        case 1: launch_bary_partial<1>(grid, st, h->lde, w, n, scale, partial); break;
        case 2: launch_bary_partial<2>(grid, st, h->lde, w, n, scale, partial); break;
        case 3: launch_bary_partial<3>(grid, st, h->lde, w, n, scale, partial); break;
        case 4: launch_bary_partial<4>(grid, st, h->lde, w, n, scale, partial); break;
        case 5: launch_bary_partial<5>(grid, st, h->lde, w, n, scale, partial); break;
        case 6: launch_bary_partial<6>(grid, st, h->lde, w, n, scale, partial); break;
        case 7: launch_bary_partial<7>(grid, st, h->lde, w, n, scale, partial); break;
        default: launch_bary_partial<8>(grid, st, h->lde, w, n, scale, partial); break;
        }
        hipLaunchKernelGGL(bary_combine_kernel, dim3((unsigned)((K * w + 255) / 256)), dim3(256), 0, st, partial,
                           chunks, (uint32_t)K, w, vfd, dout);
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipMemcpyAsync(out, dout, K * w * 24, hipMemcpyDeviceToHost, st));
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

int lvm_deep_composition(const lvm_lde *main, const lvm_lde *aux, const lvm_lde *parts, uint64_t coset_offset,
                         const uint64_t *z_points, size_t K, const uint64_t *z_parts, const uint64_t *trace_ood,
                         const uint64_t *parts_ood, const uint64_t *gammas_tr, const uint64_t *gammas_h,
                         uint64_t *dev_out) {
    if (!main || !z_points || !trace_ood || !gammas_tr || !dev_out) return fail(LVM_ERR_ARG, "null argument");
    if (K == 0 || K > (size_t)DEEP_MAX_K) return fail(LVM_ERR_ARG, "1..4 trace evaluation points");
    if ((aux && (aux->N != main->N || aux->w % 3)) || (parts && (parts->N != main->N || parts->w % 3)))
        return fail(LVM_ERR_ARG, "aux/parts tables must share the LDE domain and hold Fp3 columns");
    if (parts && (!z_parts || !parts_ood || !gammas_h)) return fail(LVM_ERR_ARG, "parts need z^P, their OOD values and gammas");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = thread_stream(d);
    const uint64_t N = main->N, wm = main->w, wa = aux ? aux->w / 3 : 0, np = parts ? parts->w / 3 : 0;
    const uint64_t cols = wm + wa;
    const uint32_t npoles = (uint32_t)K + (parts ? 1 : 0);
    // host constants: ct_k = sum_j gt_{j,k} t_j(z_k), ch = sum_j gh_j H_j(z^P)
    std::vector<uint64_t> consts(3 * (K + 1), 0), poles(3 * npoles);
    for (size_t k = 0; k < K; k++) {
        H3 acc{{0, 0, 0}};
        for (uint64_t j = 0; j < cols; j++) {
            const uint64_t *g = gammas_tr + 3 * (j * K + k), *t = trace_ood + 3 * (j * K + k);
            H3 pr = h3mul(H3{{g[0], g[1], g[2]}}, H3{{t[0], t[1], t[2]}});
            for (int c = 0; c < 3; c++) acc.c[c] = add(acc.c[c], pr.c[c]);
        }
        std::memcpy(&consts[3 * k], acc.c, 24);
        std::memcpy(&poles[3 * k], z_points + 3 * k, 24);
    }
    if (parts) {
        H3 acc{{0, 0, 0}};
        for (uint64_t j = 0; j < np; j++) {
            H3 pr = h3mul(H3{{gammas_h[3 * j], gammas_h[3 * j + 1], gammas_h[3 * j + 2]}},
                          H3{{parts_ood[3 * j], parts_ood[3 * j + 1], parts_ood[3 * j + 2]}});
            for (int c = 0; c < 3; c++) acc.c[c] = add(acc.c[c], pr.c[c]);
        }
        std::memcpy(&consts[3 * K], acc.c, 24);
        std::memcpy(&poles[3 * K], z_parts, 24);
    }
    Scratch s;
    uint64_t *dpoles, *den, *inv_den, *dconsts, *dgtr, *dgh = nullptr;
    TRYS(s.get(&dpoles, npoles * 24));
    TRYS(s.get(&den, (uint64_t)npoles * N * 24));
    TRYS(s.get(&inv_den, (uint64_t)npoles * N * 24));
    TRYS(s.get(&dconsts, (K + 1) * 24));
    TRYS(s.get(&dgtr, cols * K * 24));
    LVM_HIP(hipMemcpyAsync(dpoles, poles.data(), npoles * 24, hipMemcpyHostToDevice, st));
    LVM_HIP(hipMemcpyAsync(dconsts, consts.data(), (K + 1) * 24, hipMemcpyHostToDevice, st));
    LVM_HIP(hipMemcpyAsync(dgtr, gammas_tr, cols * K * 24, hipMemcpyHostToDevice, st));
    if (parts) {
        TRYS(s.get(&dgh, np * 24));
        LVM_HIP(hipMemcpyAsync(dgh, gammas_h, np * 24, hipMemcpyHostToDevice, st));
    }
    {
        ProfScope ps("deep_denoms", st);
        // x - z (XMinusZ, crypto/math-cuda/src/inverse.rs:241-300)
        hipLaunchKernelGGL(coset_denoms_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, den, N,
                           log2u(N), coset_offset, dpoles, npoles, 0, d->tb.fwd_lo, d->tb.fwd_hi);
        launch_batch_inverse(den, inv_den, (uint64_t)npoles * N, st);
    }
    {
        ProfScope ps("deep_rows", st);
        const uint64_t *al = aux ? aux->lde : nullptr, *pl = parts ? parts->lde : nullptr;
        switch (K) {
        case 1: launch_deep<1>(st, N, main->lde, wm, al, wa, pl, np, dgtr, dgh, dconsts, inv_den, dev_out); break;
        case 2: launch_deep<2>(st, N, main->lde, wm, al, wa, pl, np, dgtr, dgh, dconsts, inv_den, dev_out); break;
        case 3: launch_deep<3>(st, N, main->lde, wm, al, wa, pl, np, dgtr, dgh, dconsts, inv_den, dev_out); break;
        default: launch_deep<4>(st, N, main->lde, wm, al, wa, pl, np, dgtr, dgh, dconsts, inv_den, dev_out); break;
        }
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

}  // extern "C"
