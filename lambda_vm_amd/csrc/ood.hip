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

// Inverted pole factors over the LDE coset, for K poles z_k in Fp3 and the
// first `rows` physical rows p of an LDE of 2^log_N rows on offset*<w_N>
// (x_p = offset * w_N^bitrev_N(p)):
//   mode 0: out[k * rows + p] = 1 / (x_p - z_k)        (DEEP, XMinusZ)
//   mode 1: out[k * rows + p] = x_p / (z_k - x_p)      (barycentric weights)
// x - z has constant w and w^2 parts (-z1, -z2), so with a0 = x - z0 its norm
// and adjugate in Fp[w]/(w^3 - 2) are cheap base-field polynomials in a0:
//   Norm = a0^3 - 6 z1 z2 a0 - (2 z1^3 + 4 z2^3)
//   adj  = (a0^2 - 2 z1 z2, 2 z2^2 + z1 a0, z1^2 + z2 a0),  1/(x - z) = adj / Norm.
// Each thread owns E rows p = t, t + T, ... (coalesced) and inverts the
// products D(p) = prod_k Norm_k(p) of its rows with one base-field inversion
// (Montgomery's trick, prefixes kept in registers); Norm_k^-1 = D^-1 * prod_{j != k} Norm_j.
// Replaces the Fp3 batch inversion of materialised differences
// (crypto/math-cuda/src/inverse.rs:241-300, crypto/stark/src/trace.rs:900-940).
struct PoleConst {
    uint64_t z0, z1, z2, c1x3, c3, c1, c2, z1s;
};
// rows per thread E: 16 on large domains; 2 on small ones, where the chain of
// E rows (not the grid) sets the kernel's duration

template <int K, int E>
__global__ __launch_bounds__(256) void pole_inverse_kernel(uint64_t *__restrict__ out, uint64_t rows, uint32_t log_N,
                                                           uint64_t offset, const uint64_t *__restrict__ z, int mode,
                                                           const uint64_t *__restrict__ lo,
                                                           const uint64_t *__restrict__ hi) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows) return;
    PoleConst pc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint64_t z0 = z[3 * k], z1 = z[3 * k + 1], z2 = z[3 * k + 2];
        const uint64_t z12 = mul(z1, z2), c1 = dbl(z12);
        const uint64_t z1s = sqr(z1), z2s = sqr(z2);
        pc[k] = PoleConst{z0, z1, z2, add(dbl(c1), c1), add(dbl(mul(z1s, z1)), dbl(dbl(mul(z2s, z2)))), c1, dbl(z2s),
                          z1s};
    }
    auto point = [&](uint64_t p) -> uint64_t {
        const uint32_t e = log_N ? (bitrev((uint32_t)p, log_N) << (32 - log_N)) : 0u;
        return mul(offset, mul(lo[e & 0xFFFFu], hi[e >> 16]));
    };
    auto norm = [&](const PoleConst &c, uint64_t x, uint64_t &a0, uint64_t &sq) -> uint64_t {
        a0 = sub(x, c.z0);
        sq = sqr(a0);
        return sub(sub(mul(sq, a0), mul(c.c1x3, a0)), c.c3);
    };
    uint64_t pre[E];
    uint64_t acc = 1;
#pragma unroll
    for (int j = 0; j < E; j++) {
        const uint64_t p = t + (uint64_t)j * T;
        if (p < rows) {
            const uint64_t x = point(p);
            uint64_t d = 1;
#pragma unroll
            for (int k = 0; k < K; k++) {
                uint64_t a0, sq;
                d = mul(d, norm(pc[k], x, a0, sq));
            }
            acc = mul(acc, d);
        }
        pre[j] = acc;
    }
    uint64_t inv_acc = inv(acc);
#pragma unroll
    for (int j = E - 1; j >= 0; j--) {
        const uint64_t p = t + (uint64_t)j * T;
        if (p >= rows) continue;
        const uint64_t x = point(p);
        uint64_t a0[K], sq[K], nk[K];
#pragma unroll
        for (int k = 0; k < K; k++) nk[k] = norm(pc[k], x, a0[k], sq[k]);
        uint64_t d = 1;
#pragma unroll
        for (int k = 0; k < K; k++) d = mul(d, nk[k]);
        uint64_t dinv = j ? mul(inv_acc, pre[j - 1]) : inv_acc;
        inv_acc = mul(inv_acc, d);
        if (mode) dinv = mul(dinv, neg(x));
        // prefix / suffix products of the norms: Norm_k^-1 = D^-1 * pre_k * suf_k
        uint64_t sfx[K];
        sfx[K - 1] = 1;
#pragma unroll
        for (int k = K - 2; k >= 0; k--) sfx[k] = mul(sfx[k + 1], nk[k + 1]);
        uint64_t pfx = dinv;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t ni = K > 1 ? mul(pfx, sfx[k]) : pfx;
            if (k + 1 < K) pfx = mul(pfx, nk[k]);
            const PoleConst &c = pc[k];
            const fp3 adj = fp3{sub(sq[k], c.c1), add(c.c2, mul(c.z1, a0[k])), add(c.z1s, mul(c.z2, a0[k]))};
            st3(out + 3 * ((uint64_t)k * rows + p), f3_mul_base(adj, ni));
        }
    }
}

template <int E>
static void launch_pole_inverse_e(uint64_t *out, uint64_t rows, uint32_t log_N, uint64_t offset, const uint64_t *z,
                                  uint32_t K, int mode, const uint64_t *lo, const uint64_t *hi, hipStream_t st) {
    const uint64_t threads = (rows + E - 1) / E;
    const dim3 grid((unsigned)((threads + 255) / 256)), blk(256);
    switch (K) {
    case 1: hipLaunchKernelGGL((pole_inverse_kernel<1, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 2: hipLaunchKernelGGL((pole_inverse_kernel<2, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 3: hipLaunchKernelGGL((pole_inverse_kernel<3, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 4: hipLaunchKernelGGL((pole_inverse_kernel<4, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 5: hipLaunchKernelGGL((pole_inverse_kernel<5, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 6: hipLaunchKernelGGL((pole_inverse_kernel<6, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    case 7: hipLaunchKernelGGL((pole_inverse_kernel<7, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    default: hipLaunchKernelGGL((pole_inverse_kernel<8, E>), grid, blk, 0, st, out, rows, log_N, offset, z, mode, lo, hi); break;
    }
}

static void launch_pole_inverse(uint64_t *out, uint64_t rows, uint32_t log_N, uint64_t offset, const uint64_t *z,
                                uint32_t K, int mode, const uint64_t *lo, const uint64_t *hi, hipStream_t st) {
    if (!rows) return;
    if (rows >= (1u << 20))
        launch_pole_inverse_e<16>(out, rows, log_N, offset, z, K, mode, lo, hi, st);
    else
        launch_pole_inverse_e<2>(out, rows, log_N, offset, z, K, mode, lo, hi, st);
}

constexpr int BARY_MAX_K = 8;
constexpr uint64_t BARY_MAX_ROWS_PER_BLOCK = 1024;

// partial[chunk][k][c] = sum over the chunk's rows p of lde[p][c] * scale[k][p]
// block: 256 threads = 4 row lanes x 64 columns; grid (chunks, column groups),
// a chunk being `rpb` consecutive rows (fewer for small tables, so that the
// grid still has enough blocks; each thread walks rpb/4 rows, loads of four
// rows issued together). The three components of each point's sum are
// delayed-reduction dot products (one 128-bit accumulate per product, one
// reduction per thread).
template <int K>
__global__ __launch_bounds__(256) void bary_partial_kernel(const uint64_t *__restrict__ lde, uint64_t pitch,
                                                           uint64_t w, uint64_t n, uint64_t rpb,
                                                           const uint64_t *__restrict__ scale,
                                                           uint64_t *__restrict__ partial) {
    __shared__ uint64_t red[4][K][64][3];
    const int lane = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const uint64_t c = (uint64_t)blockIdx.y * 64 + lane;
    const uint64_t p0 = (uint64_t)blockIdx.x * rpb;
    acc128 acc[K][3];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) acc_zero(acc[k][j]);
    if (c < w) {
        const uint64_t pend = p0 + rpb < n ? p0 + rpb : n;
        auto row = [&](uint64_t p, uint64_t f) {
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
        };
        uint64_t p = p0 + rl;
        for (; p + 12 < pend; p += 16) {
            const uint64_t f0 = lde[p * pitch + c], f1 = lde[(p + 4) * pitch + c];
            const uint64_t f2 = lde[(p + 8) * pitch + c], f3 = lde[(p + 12) * pitch + c];
            row(p, f0);
            row(p + 4, f1);
            row(p + 8, f2);
            row(p + 12, f3);
        }
        for (; p < pend; p += 4) row(p, lde[p * pitch + c]);
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
                                uint64_t rpb, const uint64_t *scale, uint64_t *partial) {
    hipLaunchKernelGGL(bary_partial_kernel<K>, grid, dim3(256), 0, st, lde, w, w, n, rpb, scale, partial);
}

// out[k][c] = vf[k] * sum over chunks of partial[chunk][k][c]; one block per
// output (k, c): its threads sum strided chunks, then a tree in LDS.
__global__ __launch_bounds__(256) void bary_combine_kernel(const uint64_t *partial, uint64_t chunks, uint32_t K,
                                                           uint64_t w, const uint64_t *vf, uint64_t *out) {
    __shared__ uint64_t red[256][3];
    const uint64_t idx = blockIdx.x;
    const uint64_t k = idx / w, c = idx - k * w;
    const uint32_t t = threadIdx.x;
    fp3 s = fp3{0, 0, 0};
    for (uint64_t ch = t; ch < chunks; ch += 256) s = f3_add(s, ld3(partial + 3 * ((ch * K + k) * w + c)));
    red[t][0] = s.c0;
    red[t][1] = s.c1;
    red[t][2] = s.c2;
    for (uint32_t h = 128; h > 0; h >>= 1) {
        __syncthreads();
        if (t < h) {
            fp3 v = f3_add(fp3{red[t][0], red[t][1], red[t][2]}, fp3{red[t + h][0], red[t + h][1], red[t + h][2]});
            red[t][0] = v.c0;
            red[t][1] = v.c1;
            red[t][2] = v.c2;
        }
    }
    if (t == 0) st3(out + 3 * idx, f3_mul(ld3(vf + 3 * k), fp3{red[0][0], red[0][1], red[0][2]}));
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
// thread accumulates sum_j gt_{j,k} * t_j[p] as 3*KA delayed-reduction dot
// products. KA < K when the trace-term weights are geometric in k past the
// first KA - 1 points (gt_{j,k} = gt_{j,KA-1} * r^(k-KA+1) for every column j,
// which the reference's weights are: gamma^(W + j(K-1) + k - 1) for k >= 1,
// crypto/stark/src/prover.rs:2400-2460): then sum_j gt_{j,k} t_j = r^(k-KA+1)
// * sum_j gt_{j,KA-1} t_j, and rpow[k] holds r^(k-KA+1).
template <int K, int KA>
__global__ __launch_bounds__(256) void deep_kernel(uint64_t N, const uint64_t *__restrict__ main_lde, uint64_t wm,
                                                   const uint64_t *__restrict__ aux_lde, uint64_t wa,
                                                   const uint64_t *__restrict__ parts_lde, uint64_t np,
                                                   const uint64_t *__restrict__ gtr, const uint64_t *__restrict__ gh,
                                                   const uint64_t *__restrict__ consts,
                                                   const uint64_t *__restrict__ rpow,
                                                   const uint64_t *__restrict__ inv_den, uint64_t *__restrict__ out) {
    __shared__ uint64_t tile[256][DEEP_CHUNK + 1];
    __shared__ uint64_t sg[DEEP_CHUNK][KA][3];
    const int t = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * 256;
    const uint64_t p = r0 + t;
    acc128 acc[KA][3];
#pragma unroll
    for (int k = 0; k < KA; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) acc_zero(acc[k][j]);
    // the block's 256 rows x DEEP_CHUNK columns are DEEP_CHUNK values per
    // thread; the next chunk's loads are issued before the current chunk is
    // accumulated, so they are in flight during the products
    uint64_t nv[DEEP_CHUNK];
    auto fetch = [&](uint64_t j0) {
        const int cw = (int)(wm - j0 < DEEP_CHUNK ? wm - j0 : DEEP_CHUNK);
#pragma unroll
        for (int q = 0; q < DEEP_CHUNK; q++) {
            const int e = t + 256 * q, rr = e / DEEP_CHUNK, cc = e % DEEP_CHUNK;
            nv[q] = (cc < cw && r0 + rr < N) ? main_lde[(r0 + rr) * wm + j0 + cc] : 0;
        }
    };
    if (wm) fetch(0);
    for (uint64_t j0 = 0; j0 < wm; j0 += DEEP_CHUNK) {
        const int cw = (int)(wm - j0 < DEEP_CHUNK ? wm - j0 : DEEP_CHUNK);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DEEP_CHUNK; q++) {
            const int e = t + 256 * q;
            tile[e / DEEP_CHUNK][e % DEEP_CHUNK] = nv[q];
        }
        for (int e = t; e < DEEP_CHUNK * KA * 3; e += 256) {
            const int cc = e / (KA * 3), rem = e % (KA * 3);
            sg[cc][rem / 3][rem % 3] = cc < cw ? gtr[3 * ((j0 + cc) * K + rem / 3) + rem % 3] : 0;
        }
        __syncthreads();
        if (j0 + DEEP_CHUNK < wm) fetch(j0 + DEEP_CHUNK);
#pragma unroll 4
        for (int cc = 0; cc < DEEP_CHUNK; cc++) {
            const uint64_t v = tile[t][cc];
#pragma unroll
            for (int k = 0; k < KA; k++) {
                acc_mad(acc[k][0], sg[cc][k][0], v);
                acc_mad(acc[k][1], sg[cc][k][1], v);
                acc_mad(acc[k][2], sg[cc][k][2], v);
            }
        }
    }
    if (p >= N) return;
    fp3 tk[K];
#pragma unroll
    for (int k = 0; k < KA; k++) tk[k] = fp3{acc_reduce(acc[k][0]), acc_reduce(acc[k][1]), acc_reduce(acc[k][2])};
    if (aux_lde) {
        const uint64_t *arow = aux_lde + p * 3 * wa;
        for (uint64_t j = 0; j < wa; j++) {
            fp3 v = ld3(arow + 3 * j);
#pragma unroll
            for (int k = 0; k < KA; k++) tk[k] = f3_add(tk[k], f3_mul(ld3(gtr + 3 * ((wm + j) * K + k)), v));
        }
    }
#pragma unroll
    for (int k = KA; k < K; k++) tk[k] = f3_mul(ld3(rpow + 3 * k), tk[KA - 1]);
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

template <int K, int KA>
static void launch_deep(hipStream_t st, uint64_t N, const uint64_t *main_lde, uint64_t wm, const uint64_t *aux_lde,
                        uint64_t wa, const uint64_t *parts_lde, uint64_t np, const uint64_t *gtr, const uint64_t *gh,
                        const uint64_t *consts, const uint64_t *rpow, const uint64_t *inv_den, uint64_t *out) {
    hipLaunchKernelGGL((deep_kernel<K, KA>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, main_lde, wm,
                       aux_lde, wa, parts_lde, np, gtr, gh, consts, rpow, inv_den, out);
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
    uint64_t *dz, *scale, *partial, *vfd, *dout;
    // rows per block: 1024, fewer (down to 64) when the grid would have under
    // ~1024 blocks
    const uint64_t cgroups = (w + 63) / 64;
    uint64_t rpb = BARY_MAX_ROWS_PER_BLOCK;
    while (rpb > 64 && ((n + rpb - 1) / rpb) * cgroups < 1024) rpb /= 2;
    const uint64_t chunks = (n + rpb - 1) / rpb;
    TRYS(s.get(&dz, K * 24));
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
    StageScope ss(st);
    LVM_HIP(stage_upload(dz, z, K * 24, st));
    LVM_HIP(stage_upload(vfd, vf.data(), K * 24, st));
    {
        ProfScope ps("bary_denoms", st);
        // scale[k][p] = x_p / (z_k - x_p)
        launch_pole_inverse(scale, n, logN, coset_offset, dz, (uint32_t)K, 1, d->tb.fwd_lo, d->tb.fwd_hi, st);
    }
    {
        ProfScope ps("bary_sum", st);
        dim3 grid((unsigned)chunks, (unsigned)((w + 63) / 64));
        switch (K) {
        case 1: launch_bary_partial<1>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 2: launch_bary_partial<2>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 3: launch_bary_partial<3>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 4: launch_bary_partial<4>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 5: launch_bary_partial<5>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 6: launch_bary_partial<6>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        case 7: launch_bary_partial<7>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        default: launch_bary_partial<8>(grid, st, h->lde, w, n, rpb, scale, partial); break;
        }
        hipLaunchKernelGGL(bary_combine_kernel, dim3((unsigned)(K * w)), dim3(256), 0, st, partial, chunks,
                           (uint32_t)K, w, vfd, dout);
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(stage_download(out, dout, K * w * 24, st));
    LVM_HIP(stage_sync(st));
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
    // geometric trace-term weights past k = 1: gt_{j,k} = gt_{j,k-1} * r for all j, k >= 2
    std::vector<uint64_t> rp(3 * K, 0);
    bool geo = K > 2;
    if (geo) {
        const H3 g01{{gammas_tr[3], gammas_tr[4], gammas_tr[5]}}, g02{{gammas_tr[6], gammas_tr[7], gammas_tr[8]}};
        if (canon(g01.c[0]) == 0 && canon(g01.c[1]) == 0 && canon(g01.c[2]) == 0) {
            geo = false;
        } else {
            const fp3 ri = f3_inv(fp3{g01.c[0], g01.c[1], g01.c[2]});
            const H3 r = h3mul(g02, H3{{ri.c0, ri.c1, ri.c2}});
            H3 pw{{1, 0, 0}};
            for (size_t k = 2; k < K; k++) {
                pw = h3mul(pw, r);
                std::memcpy(&rp[3 * k], pw.c, 24);
            }
            for (uint64_t j = 0; j < cols && geo; j++)
                for (size_t k = 2; k < K && geo; k++) {
                    const uint64_t *a = gammas_tr + 3 * (j * K + k - 1), *b = gammas_tr + 3 * (j * K + k);
                    const H3 want = h3mul(H3{{a[0], a[1], a[2]}}, r);
                    for (int c = 0; c < 3; c++) geo = geo && canon(want.c[c]) == canon(b[c]);
                }
        }
    }
    Scratch s;
    uint64_t *dpoles, *inv_den, *dconsts, *dgtr, *drp, *dgh = nullptr;
    TRYS(s.get(&dpoles, npoles * 24));
    TRYS(s.get(&inv_den, (uint64_t)npoles * N * 24));
    TRYS(s.get(&dconsts, (K + 1) * 24));
    TRYS(s.get(&dgtr, cols * K * 24));
    TRYS(s.get(&drp, K * 24));
    StageScope ss(st);
    LVM_HIP(stage_upload(dpoles, poles.data(), npoles * 24, st));
    LVM_HIP(stage_upload(dconsts, consts.data(), (K + 1) * 24, st));
    LVM_HIP(stage_upload(dgtr, gammas_tr, cols * K * 24, st));
    LVM_HIP(stage_upload(drp, rp.data(), K * 24, st));
    if (parts) {
        TRYS(s.get(&dgh, np * 24));
        LVM_HIP(stage_upload(dgh, gammas_h, np * 24, st));
    }
    {
        ProfScope ps("deep_denoms", st);
        // 1 / (x - z) (XMinusZ, crypto/math-cuda/src/inverse.rs:241-300)
        launch_pole_inverse(inv_den, N, log2u(N), coset_offset, dpoles, npoles, 0, d->tb.fwd_lo, d->tb.fwd_hi, st);
    }
    {
        ProfScope ps("deep_rows", st);
        const uint64_t *al = aux ? aux->lde : nullptr, *pl = parts ? parts->lde : nullptr;
        const uint64_t *m = main->lde;
        switch (K * 2 + (geo ? 1 : 0)) {
        case 2: launch_deep<1, 1>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        case 4: launch_deep<2, 2>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        case 6: launch_deep<3, 3>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        case 7: launch_deep<3, 2>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        case 8: launch_deep<4, 4>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        default: launch_deep<4, 2>(st, N, m, wm, al, wa, pl, np, dgtr, dgh, dconsts, drp, inv_den, dev_out); break;
        }
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(stage_sync(st));
    return LVM_OK;
}

}  // extern "C"
