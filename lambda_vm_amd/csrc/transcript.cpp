// Host-side Fiat-Shamir transcript and the FRI commit-phase driver.
//
// The transcript follows crypto/crypto/src/fiat_shamir/default_transcript.rs:
// a Keccak-256 sponge over absorbed bytes; sample() = finalize, reverse the
// 32 bytes, restart the sponge with them (:75-81); field and u64 samples read
// 8 buffered bytes big-endian (:86-95) with rejection of values >= p
// (goldilocks.rs:548-555) or below 2^64 mod ub (:136-145); appends
// invalidate the output buffer (:113-126). It lives on the host because
// every challenge depends on a root the host has just received, exactly as
// in the reference prover.
//
// lvm_fri_commit_phase drives crypto/stark/src/fri/mod.rs:27-151 over a
// device-resident codeword: per committed layer zeta <- transcript, fold +
// leaves + tree on the device, root -> transcript; then the final fold, the
// terminal coefficients (crypto/stark/src/fri/terminal.rs:72-103) and their
// absorption.
#include <cstring>
#include <string>
#include <vector>

#include "lvm_internal.hpp"

namespace lvm {

// ---- host Keccak-f[1600] (FIPS 202) ----
static const uint64_t HRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
static const int PI_SRC[25] = {0, 6, 12, 18, 24, 3, 9, 10, 16, 22, 1, 7, 13, 19, 20, 4, 5, 11, 17, 23, 2, 8, 14, 15, 21};
static const int PI_ROT[25] = {0, 44, 43, 21, 14, 28, 20, 3, 45, 61, 1, 6, 25, 8, 18, 27, 36, 10, 15, 56, 62, 55, 39, 41, 2};

static inline uint64_t rol(uint64_t v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }

void host_keccak_f(uint64_t s[25]) {
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
        for (int x = 0; x < 5; x++) {
            uint64_t D = C[(x + 4) % 5] ^ rol(C[(x + 1) % 5], 1);
            for (int y = 0; y < 25; y += 5) s[y + x] ^= D;
        }
        // B[dst] = rot(A[PI_SRC[dst]], PI_ROT[dst])
        for (int i = 0; i < 25; i++) B[i] = rol(s[PI_SRC[i]], PI_ROT[i]);
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; x++) s[y + x] = B[y + x] ^ (~B[y + (x + 1) % 5] & B[y + (x + 2) % 5]);
        s[0] ^= HRC[round];
    }
}

struct Sponge {
    uint64_t st[25];
    size_t pos;
    void reset() {
        std::memset(st, 0, sizeof(st));
        pos = 0;
    }
    void absorb(const uint8_t *p, size_t n) {
        uint8_t *b = (uint8_t *)st;
        while (n) {
            size_t take = 136 - pos;
            if (take > n) take = n;
            for (size_t i = 0; i < take; i++) b[pos + i] ^= p[i];
            pos += take;
            p += take;
            n -= take;
            if (pos == 136) {
                host_keccak_f(st);
                pos = 0;
            }
        }
    }
    void digest(uint8_t out[32]) const {
        Sponge c = *this;
        uint8_t *b = (uint8_t *)c.st;
        b[c.pos] ^= 0x01;
        b[135] ^= 0x80;
        host_keccak_f(c.st);
        std::memcpy(out, c.st, 32);
    }
};

void host_keccak256(const uint8_t *p, size_t n, uint8_t out[32]) {
    Sponge s;
    s.reset();
    s.absorb(p, n);
    s.digest(out);
}

}  // namespace lvm

using namespace lvm;

struct lvm_transcript {
    Sponge h;
    uint8_t out_buf[32];
    size_t out_pos;
};

namespace {
void be64(uint64_t v, uint8_t *o) {
    for (int i = 0; i < 8; i++) o[i] = (uint8_t)(v >> (56 - 8 * i));
}

uint64_t next_u64(lvm_transcript *t);
}  // namespace

extern "C" {

int lvm_transcript_new(const uint8_t *data, size_t len, lvm_transcript **out) {
    if (!out || (len && !data)) return fail(LVM_ERR_ARG, "null argument");
    lvm_transcript *t = new lvm_transcript();
    t->h.reset();
    t->out_pos = 32;
    if (len) t->h.absorb(data, len);
    *out = t;
    return LVM_OK;
}

int lvm_transcript_clone(const lvm_transcript *t, lvm_transcript **out) {
    if (!t || !out) return fail(LVM_ERR_ARG, "null argument");
    *out = new lvm_transcript(*t);
    return LVM_OK;
}

int lvm_transcript_release(lvm_transcript *t) {
    delete t;
    return LVM_OK;
}

int lvm_transcript_append_bytes(lvm_transcript *t, const uint8_t *data, size_t len) {
    if (!t || (len && !data)) return fail(LVM_ERR_ARG, "null argument");
    t->out_pos = 32;
    t->h.absorb(data, len);
    return LVM_OK;
}

int lvm_transcript_append_base(lvm_transcript *t, uint64_t v) {
    if (!t) return fail(LVM_ERR_ARG, "null argument");
    uint8_t b[8];
    be64(canon(v), b);
    return lvm_transcript_append_bytes(t, b, 8);
}

int lvm_transcript_append_fp3(lvm_transcript *t, const uint64_t v[3]) {
    if (!t || !v) return fail(LVM_ERR_ARG, "null argument");
    uint8_t b[24];
    for (int i = 0; i < 3; i++) be64(canon(v[i]), b + 8 * i);
    return lvm_transcript_append_bytes(t, b, 24);
}

int lvm_transcript_append_fp3_many(lvm_transcript *t, const uint64_t *v, size_t count) {
    if (!t || (!v && count)) return fail(LVM_ERR_ARG, "null argument");
    std::vector<uint8_t> b(24 * count);
    for (size_t k = 0; k < 3 * count; k++) be64(canon(v[k]), b.data() + 8 * k);
    return lvm_transcript_append_bytes(t, b.data(), b.size());
}

int lvm_fp3_powers(const uint64_t base[3], size_t count, uint64_t *out) {
    if (!base || (!out && count)) return fail(LVM_ERR_ARG, "null argument");
    const fp3 g{base[0], base[1], base[2]};
    fp3 cur{1, 0, 0};
    for (size_t i = 0; i < count; i++) {
        out[3 * i] = canon(cur.c0);
        out[3 * i + 1] = canon(cur.c1);
        out[3 * i + 2] = canon(cur.c2);
        cur = f3_mul(cur, g);
    }
    return LVM_OK;
}

int lvm_transcript_sample(lvm_transcript *t, uint8_t out[32]) {
    if (!t || !out) return fail(LVM_ERR_ARG, "null argument");
    uint8_t d[32];
    t->h.digest(d);
    for (int i = 0; i < 32; i++) out[i] = d[31 - i];
    t->h.reset();
    t->h.absorb(out, 32);
    t->out_pos = 32;
    return LVM_OK;
}

int lvm_transcript_state(const lvm_transcript *t, uint8_t out[32]) {
    if (!t || !out) return fail(LVM_ERR_ARG, "null argument");
    t->h.digest(out);
    return LVM_OK;
}

int lvm_transcript_sample_base(lvm_transcript *t, uint64_t *out) {
    if (!t || !out) return fail(LVM_ERR_ARG, "null argument");
    for (;;) {
        uint64_t c = next_u64(t);
        if (c < P) {
            *out = c;
            return LVM_OK;
        }
    }
}

int lvm_transcript_sample_fp3(lvm_transcript *t, uint64_t out[3]) {
    if (!t || !out) return fail(LVM_ERR_ARG, "null argument");
    for (int i = 0; i < 3; i++) lvm_transcript_sample_base(t, &out[i]);
    return LVM_OK;
}

int lvm_transcript_sample_u64(lvm_transcript *t, uint64_t upper_bound, uint64_t *out) {
    if (!t || !out || upper_bound == 0) return fail(LVM_ERR_ARG, "bad argument");
    uint64_t threshold = (0 - upper_bound) % upper_bound;
    for (;;) {
        uint64_t c = next_u64(t);
        if (c >= threshold) {
            *out = c % upper_bound;
            return LVM_OK;
        }
    }
}

}  // extern "C"

namespace {
uint64_t next_u64(lvm_transcript *t) {
    if (t->out_pos + 8 > 32) {
        uint8_t s[32];
        lvm_transcript_sample(t, s);
        std::memcpy(t->out_buf, s, 32);
        t->out_pos = 0;
    }
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v = (v << 8) | t->out_buf[t->out_pos + i];
    t->out_pos += 8;
    return v;
}

// ---- host Fp3 helpers for the terminal polynomial ----
struct F3 {
    uint64_t c[3];
};
F3 f3add(F3 a, F3 b) { return F3{{add(a.c[0], b.c[0]), add(a.c[1], b.c[1]), add(a.c[2], b.c[2])}}; }
F3 f3subh(F3 a, F3 b) { return F3{{sub(a.c[0], b.c[0]), sub(a.c[1], b.c[1]), sub(a.c[2], b.c[2])}}; }
F3 f3scale(F3 a, uint64_t s) { return F3{{mul(s, a.c[0]), mul(s, a.c[1]), mul(s, a.c[2])}}; }

uint64_t root_of_unity(unsigned order) {
    uint64_t r = 1753635133440165772ULL;  // goldilocks.rs:540
    for (unsigned i = 0; i < 32 - order; i++) r = sqr(r);
    return order ? r : 1;
}

uint32_t brev_h(uint32_t i, unsigned bits) {
    uint32_t r = 0;
    for (unsigned b = 0; b < bits; b++) r |= ((i >> b) & 1u) << (bits - 1 - b);
    return r;
}

// Coefficients of the degree < 2^k polynomial whose evaluations on the 2^k
// sub-coset offset*<w^blowup> are every blowup-th natural-order entry of the
// bit-reversed terminal codeword (crypto/stark/src/fri/terminal.rs:72-103):
// inverse DFT, then scale coefficient i by 2^-k * offset^-i.
void terminal_coeffs(const std::vector<F3> &cw_bitrev, uint64_t offset, unsigned k, std::vector<F3> &coeffs) {
    const size_t len = cw_bitrev.size(), keep = (size_t)1 << k, blowup = len / keep;
    const unsigned ll = (unsigned)__builtin_ctzll(len);
    std::vector<F3> v(keep);
    for (size_t m = 0; m < keep; m++) v[m] = cw_bitrev[brev_h((uint32_t)(m * blowup), ll)];
    // iterative radix-2 DIT inverse DFT on bit-reversed input
    for (size_t i = 0; i < keep; i++) {
        size_t j = brev_h((uint32_t)i, k);
        if (j > i) std::swap(v[i], v[j]);
    }
    for (unsigned s = 1; s <= k; s++) {
        size_t lenb = (size_t)1 << s, h = lenb / 2;
        uint64_t wl = inv(root_of_unity(s));
        for (size_t b = 0; b < keep; b += lenb) {
            uint64_t w = 1;
            for (size_t j = 0; j < h; j++) {
                F3 a = v[b + j], t = f3scale(v[b + j + h], w);
                v[b + j] = f3add(a, t);
                v[b + j + h] = f3subh(a, t);
                w = mul(w, wl);
            }
        }
    }
    uint64_t pw = inv((uint64_t)keep), oinv = inv(offset);
    coeffs.resize(keep);
    for (size_t i = 0; i < keep; i++) {
        coeffs[i] = f3scale(v[i], pw);
        pw = mul(pw, oinv);
    }
}
}  // namespace

extern "C" {

int lvm_fri_commit_phase(lvm_fri *f, lvm_transcript *t, unsigned blowup_log, unsigned k, uint8_t *roots_out,
                         size_t *num_committed_out, uint64_t *coeffs_out, size_t *num_coeffs_out) {
    if (!f || !t) return fail(LVM_ERR_ARG, "null argument");
    size_t len = 0;
    uint64_t offset = 0;
    int rc = lvm_fri_current(f, nullptr, &len, &offset);
    if (rc) return rc;
    const unsigned lde_log = (unsigned)__builtin_ctzll(len);
    // FriFoldLayout (crypto/stark/src/fri/terminal.rs:45-55)
    unsigned terminal_log = blowup_log + k < lde_log ? blowup_log + k : lde_log;
    if (terminal_log < blowup_log) return fail(LVM_ERR_ARG, "codeword shorter than the blowup");
    const unsigned total_folds = lde_log - terminal_log;
    const size_t num_committed = total_folds ? total_folds - 1 : 0;
    const unsigned eff_k = terminal_log - blowup_log;
    // the transcript is restored if the device fails mid-loop (gpu_lde.rs:3196-3229)
    lvm_transcript snapshot = *t;
    for (size_t l = 0; l < num_committed; l++) {
        uint64_t zeta[3];
        lvm_transcript_sample_fp3(t, zeta);
        uint8_t root[32];
        rc = lvm_fri_fold_and_commit_layer(f, zeta, root);
        if (rc) {
            *t = snapshot;
            return rc;
        }
        if (roots_out) std::memcpy(roots_out + 32 * l, root, 32);
        lvm_transcript_append_bytes(t, root, 32);
    }
    if (total_folds > 0) {
        uint64_t zeta[3];
        lvm_transcript_sample_fp3(t, zeta);
        rc = lvm_fri_fold(f, zeta);
        if (rc) {
            *t = snapshot;
            return rc;
        }
    }
    rc = lvm_fri_current(f, nullptr, &len, &offset);
    if (rc) {
        *t = snapshot;
        return rc;
    }
    std::vector<F3> cw(len);
    rc = lvm_fri_current(f, (uint64_t *)cw.data(), &len, &offset);
    if (rc) {
        *t = snapshot;
        return rc;
    }
    std::vector<F3> coeffs;
    terminal_coeffs(cw, offset, eff_k, coeffs);
    for (auto &c : coeffs) lvm_transcript_append_fp3(t, c.c);
    if (coeffs_out) std::memcpy(coeffs_out, coeffs.data(), coeffs.size() * 24);
    if (num_committed_out) *num_committed_out = num_committed;
    if (num_coeffs_out) *num_coeffs_out = coeffs.size();
    return LVM_OK;
}

int lvm_fri_commit_phase_batch(lvm_fri *const *fs, lvm_transcript *const *ts, size_t count,
                               const unsigned *blowup_log, unsigned k, uint8_t *const *roots_out,
                               size_t *num_committed_out, uint64_t *const *coeffs_out, size_t *num_coeffs_out) {
    if (count && (!fs || !ts || !blowup_log)) return fail(LVM_ERR_ARG, "null argument");
    struct Plan {
        size_t committed;
        unsigned folds, eff_k;
    };
    std::vector<Plan> plan(count);
    size_t max_committed = 0;
    for (size_t t = 0; t < count; t++) {
        if (!fs[t] || !ts[t]) return fail(LVM_ERR_ARG, "null FRI state or transcript");
        size_t len = 0;
        uint64_t offset = 0;
        int rc = lvm_fri_current(fs[t], nullptr, &len, &offset);
        if (rc) return rc;
        const unsigned lde_log = (unsigned)__builtin_ctzll(len);
        // FriFoldLayout (crypto/stark/src/fri/terminal.rs:45-55)
        const unsigned terminal_log = blowup_log[t] + k < lde_log ? blowup_log[t] + k : lde_log;
        if (terminal_log < blowup_log[t]) return fail(LVM_ERR_ARG, "codeword shorter than the blowup");
        plan[t].folds = lde_log - terminal_log;
        plan[t].committed = plan[t].folds ? plan[t].folds - 1 : 0;
        plan[t].eff_k = terminal_log - blowup_log[t];
        max_committed = plan[t].committed > max_committed ? plan[t].committed : max_committed;
    }
    // every transcript is restored if the device fails mid-loop (gpu_lde.rs:3196-3229)
    std::vector<lvm_transcript> snapshot;
    snapshot.reserve(count);
    for (size_t t = 0; t < count; t++) snapshot.push_back(*ts[t]);
    auto restore = [&](int rc) {
        for (size_t t = 0; t < count; t++) *ts[t] = snapshot[t];
        return rc;
    };
    std::vector<lvm_fri *> act;
    std::vector<size_t> idx;
    std::vector<uint64_t> zetas;
    std::vector<uint8_t> roots;
    for (size_t l = 0; l <= max_committed; l++) {
        // depth l < committed: a committed fold; l == committed: the final fold
        for (int final_fold = 0; final_fold < 2; final_fold++) {
            act.clear();
            idx.clear();
            zetas.clear();
            for (size_t t = 0; t < count; t++) {
                const bool take = final_fold ? (plan[t].folds > 0 && l == plan[t].committed) : l < plan[t].committed;
                if (!take) continue;
                uint64_t z[3];
                lvm_transcript_sample_fp3(ts[t], z);
                zetas.insert(zetas.end(), z, z + 3);
                act.push_back(fs[t]);
                idx.push_back(t);
            }
            if (act.empty()) continue;
            roots.assign(act.size() * 32, 0);
            int rc = fri_fold_batch(act.data(), act.size(), zetas.data(), !final_fold, roots.data());
            if (rc) return restore(rc);
            if (final_fold) continue;
            for (size_t i = 0; i < act.size(); i++) {
                const size_t t = idx[i];
                if (roots_out && roots_out[t]) std::memcpy(roots_out[t] + 32 * l, roots.data() + 32 * i, 32);
                lvm_transcript_append_bytes(ts[t], roots.data() + 32 * i, 32);
            }
        }
    }
    for (size_t t = 0; t < count; t++) {
        size_t len = 0;
        uint64_t offset = 0;
        int rc = lvm_fri_current(fs[t], nullptr, &len, &offset);
        std::vector<F3> cw(len);
        if (!rc) rc = lvm_fri_current(fs[t], (uint64_t *)cw.data(), &len, &offset);
        if (rc) return restore(rc);
        std::vector<F3> coeffs;
        terminal_coeffs(cw, offset, plan[t].eff_k, coeffs);
        for (auto &c : coeffs) lvm_transcript_append_fp3(ts[t], c.c);
        if (coeffs_out && coeffs_out[t]) std::memcpy(coeffs_out[t], coeffs.data(), coeffs.size() * 24);
        if (num_committed_out) num_committed_out[t] = plan[t].committed;
        if (num_coeffs_out) num_coeffs_out[t] = coeffs.size();
    }
    return LVM_OK;
}

}  // extern "C"
