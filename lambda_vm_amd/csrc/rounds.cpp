// Rounds 3-4 of a committed table in one call: the whole sequence of
// library calls the host would otherwise drive one by one, so that a table's
// remaining rounds cost no interpreter time between the device steps.
//   R3  z <- transcript, resampled while z^n = 1 or z^N = g^N (z in the trace
//       domain or the LDE coset; crypto/crypto/src/fiat_shamir/is_transcript.rs
//       :44-75); frame points z, z*w_n (crypto/stark/src/trace.rs:1112-1132);
//       barycentric values of every column at both points, appended point-major
//       (crypto/stark/src/prover.rs:4547-4560)
//   R4  gamma <- transcript; trace-term weight of column j at point k is
//       gamma^j (k = 0) or gamma^(m + j(K-1) + k - 1) (crypto/stark/src/ood.rs
//       :57-92); DEEP codeword; FRI commit phase (fri/mod.rs:27-151); query
//       indexes sample_u64(N/2) (prover.rs:2324-2333); FRI query phase and the
//       trace openings.
#include <cstring>
#include <vector>

#include "lvm_internal.hpp"

using namespace lvm;

namespace {
bool f3_is(const fp3 &a, uint64_t c0) { return canon(a.c0) == canon(c0) && canon(a.c1) == 0 && canon(a.c2) == 0; }
bool is_pow2_size(size_t x) { return x && !(x & (x - 1)); }
}  // namespace

extern "C" {

int lvm_prove_committed_table(const lvm_lde *h, lvm_transcript *t, size_t n, unsigned blowup, uint64_t offset,
                              unsigned fri_k, size_t num_queries, lvm_table_rounds *r) {
    if (!h || !t || !r || !r->ood || !r->fri_roots || !r->fri_coeffs || !r->iotas || !r->trace_paths ||
        !r->trace_rows)
        return fail(LVM_ERR_ARG, "null argument");
    if (!is_pow2_size(n) || !is_pow2_size(blowup) || n < 2) return fail(LVM_ERR_ARG, "n and blowup must be powers of two");
    size_t lde_rows = 0, words = 0, trace_rows = 0;
    int rc = lvm_lde_shape(h, &lde_rows, &words, &trace_rows);
    if (rc) return rc;
    const size_t N = n * blowup;
    if (lde_rows != N || trace_rows != n) return fail(LVM_ERR_ARG, "LDE shape differs from n x blowup");
    const size_t m = words;
    constexpr size_t K = 2;
    // R3: out-of-domain point
    const uint64_t g_pow = pow(offset, N);
    fp3 z{0, 0, 0};
    for (;;) {
        uint64_t s[3];
        rc = lvm_transcript_sample_fp3(t, s);
        if (rc) return rc;
        z = fp3{s[0], s[1], s[2]};
        const fp3 zt = f3_pow(z, n);
        if (f3_is(zt, 1)) continue;
        if (f3_is(f3_pow(zt, blowup), g_pow)) continue;
        break;
    }
    const fp3 z1 = f3_mul_base(z, root_of_unity((unsigned)__builtin_ctzll(n)));
    const uint64_t pts[6] = {canon(z.c0), canon(z.c1), canon(z.c2), canon(z1.c0), canon(z1.c1), canon(z1.c2)};
    std::memcpy(r->z, pts, 24);
    rc = lvm_barycentric_base_multi(h, offset, pts, K, r->ood);  // (K, m, 3)
    if (rc) return rc;
    rc = lvm_transcript_append_fp3_many(t, r->ood, K * m);
    if (rc) return rc;
    // R4: DEEP weights, codeword, FRI
    uint64_t gamma[3];
    rc = lvm_transcript_sample_fp3(t, gamma);
    if (rc) return rc;
    std::vector<uint64_t> gp(3 * m * K), gammas(3 * m * K), trace_ood(3 * m * K);
    rc = lvm_fp3_powers(gamma, m * K, gp.data());
    if (rc) return rc;
    for (size_t j = 0; j < m; j++)
        for (size_t k = 0; k < K; k++) {
            const size_t gi = k == 0 ? j : m + j * (K - 1) + k - 1;
            std::memcpy(&gammas[3 * (j * K + k)], &gp[3 * gi], 24);
            std::memcpy(&trace_ood[3 * (j * K + k)], r->ood + 3 * (k * m + j), 24);
        }
    void *deep = nullptr;
    rc = lvm_dev_alloc(&deep, N * 24);
    if (rc) return rc;
    rc = lvm_deep_composition(h, nullptr, nullptr, offset, pts, K, nullptr, trace_ood.data(), nullptr, gammas.data(),
                              nullptr, (uint64_t *)deep);
    if (rc) {
        lvm_dev_free(deep);
        return rc;
    }
    lvm_fri *f = nullptr;
    rc = lvm_fri_new((const uint64_t *)deep, N, offset, 2, &f);  // the FRI state takes the buffer over
    if (rc) {
        lvm_dev_free(deep);
        return rc;
    }
    const unsigned blowup_log = (unsigned)__builtin_ctzll(blowup);
    rc = lvm_fri_commit_phase(f, t, blowup_log, fri_k, r->fri_roots, &r->num_fri_roots, r->fri_coeffs,
                              &r->num_fri_coeffs);
    for (size_t q = 0; rc == LVM_OK && q < num_queries; q++) rc = lvm_transcript_sample_u64(t, N / 2, &r->iotas[q]);
    if (rc == LVM_OK && num_queries) rc = lvm_fri_query(f, r->iotas, num_queries, r->fri_sym, r->fri_paths);
    if (rc == LVM_OK && num_queries) rc = lvm_lde_open(h, r->iotas, num_queries, r->trace_paths, r->trace_rows);
    lvm_fri_release(f);
    return rc;
}

}  // extern "C"
