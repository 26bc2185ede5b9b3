// Tree packing of the column LWEs (respond_one_packing, src/pir/respond.rs:
// 213-268; pack_lwes / pack_lwes_inner / homomorphic_automorph,
// src/inspiring/automorph_pack.rs:44-78, 171-256, 418-487) on gfx950.
//
// The reference pads the m column LWEs to d leaves and recurses over all of
// them.  A subtree whose leaves are all padding is exactly zero (every step
// is linear in its inputs), so only the nodes that hold a column take work:
// at level l (1..log d) the nodes are s < min(m, 2^(log d - l)), node s
// combines node s of level l-1 (even) with node s + 2^(log d - l) (odd, when
// that index is < m).  Skipping the zero nodes leaves every output word as
// the full recursion computes it.
//
// Node state: a in the coefficient domain (the gadget digits of the
// automorphism need it) and in the NTT domain (a^), b in the NTT domain (b
// only ever enters linearly: the automorphism is a slot gather and y =
// X^(d/2^l) a pointwise product), all canonical u32 residues, [3][K][N] per
// node.  One workgroup of K x N/4 threads works on one node, the k-th group
// of N/4 threads (four values each, kTreeV) on residue k.
//
// One level (pack_lwes_inner with t = 2^l + 1 and the galois key of index
// log d - l, whose rows are kept as NTT-domain Montgomery words):
//   s0 = even + y odd,  s1 = even - y odd          (y odd: negacyclic shift)
//   D_i = NTT(G^-1(tau_t(s1.a)))                    (gadget digits, i < ell)
//   a^  = s0.a^ + sum_i D_i * key[i].a,  a = s0.a + INTT(sum_i D_i * key[i].a)
//   b   = s0.b + tau_t(s1.b) + sum_i D_i * key[i].b (NTT domain)
// which is s0 + homomorphic_automorph(s1, t) term for term.  When the digits
// cover the composite value (q <= z^ell), sum_i z^i D_i = tau_t(s1.a)^ exactly
// and the last digit's transform is recovered from the others instead of
// computed: D_last = z^-(ell-1) (tau_t(s1.a^) - sum_{i<ell-1} z^i D_i), with
// tau_t(s1.a^) a slot gather.
//
// k_tree_leaves computes each column's ciphertext from the shard digits and
// the query's RGSW rows (the first dimension), takes the LWE of its
// coefficient 0 (sample_extract_coeff0: a stays the column's a, b is
// coefficient 0 of the column's b, n^-1 times the sum of its NTT slots) and
// runs the levels at which no node has an odd child.  k_tree_combine runs one
// later level; at the root it finishes pack_lwes: b leaves the NTT domain and
// b_val * d is added at coefficient z (automorph_pack.rs:473-484).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "reg_ntt.hpp"
#include "tree.hpp"

namespace inspire {

namespace {

// Values per thread of the register-resident transforms: four (N/4 threads
// per residue, four waves per SIMD at d = 2048) rather than eight.
constexpr int kTreeV = 4;

template <int ELL, bool REC>
struct TreeDigits {
    static constexpr int ND = REC ? ELL - 1 : ELL;  // digits transformed
    static constexpr int DB = ND > 0 ? ND : 1;      // LDS exchange buffers per residue
};

template <int LOGN, int K, int ELL, bool REC>
struct TreeShape {
    static constexpr int N = 1 << LOGN, T = N / kTreeV, LP = LayV<LOGN, kTreeV>::LP;
    static constexpr size_t kTw = (size_t)K * N;          // forward Montgomery twiddles (the inverse reads them mirrored)
    static constexpr size_t kA = (size_t)K * LP;          // s1.a (automorphism source), then INTT exchange
    static constexpr size_t kB = (size_t)K * LP;          // s1.b (automorphism source)
    static constexpr size_t kH = REC ? (size_t)K * LP : 0;  // s1.a^ (automorphism source of the recovered digit)
    static constexpr size_t kD = (size_t)K * TreeDigits<ELL, REC>::DB * LP;  // digit-transform exchange
    static constexpr int W = T < 64 ? T : 64;              // lanes of one reduction group
    static constexpr size_t kRed = 2 * K * (T / W);       // u64 partial sums (two words each)
    static constexpr size_t kBytes = 4 * (kTw + kA + kB + kH + kD + kRed);
};

template <int LOGN, int K, int ELL, bool REC>
struct TreeLds {
    using S = TreeShape<LOGN, K, ELL, REC>;
    uint32_t *twf, *abuf, *bbuf, *hbuf, *dbuf;
    uint64_t* red;
    __device__ explicit TreeLds(uint32_t* lds) {
        twf = lds;
        abuf = twf + K * S::N;
        bbuf = abuf + S::kA;
        hbuf = bbuf + S::kB;
        dbuf = hbuf + S::kH;
        red = reinterpret_cast<uint64_t*>(dbuf + S::kD);
    }
};

// Forward twiddles of every residue into LDS, Montgomery form.
template <int LOGN, int K, int ELL, bool REC>
__device__ __forceinline__ void stage_twiddles(const TreeLds<LOGN, K, ELL, REC>& L, const CtxDev& C) {
    constexpr int N = 1 << LOGN;
    for (int e = threadIdx.x; e < K * N; e += blockDim.x) {
        const int kk = e / N, i = e - kk * N;
        const ModDev& D = C.m[kk];
        L.twf[e] = csub(shoup_mul(D.tw[i], C.mont[kk], C.montp[kk], D.q), D.q);
    }
}

__device__ __forceinline__ uint32_t addq(uint32_t x, uint32_t y, uint32_t q) { return csub(x + y, q); }
__device__ __forceinline__ uint32_t subq(uint32_t x, uint32_t y, uint32_t q) { return csub(x + q - y, q); }
__device__ __forceinline__ uint32_t negq(uint32_t x, uint32_t q) { return x ? q - x : 0u; }

// Inverse of an odd t modulo 2^32 (Newton: each step doubles the correct bits).
__device__ __forceinline__ uint32_t inv_odd(uint32_t t) {
    uint32_t y = t;  // correct to 3 bits
#pragma unroll
    for (int i = 0; i < 4; i++) y *= 2u - t * y;
    return y;
}

// Slot p of NTT(tau_t(x)) is slot br((t (2 br(p) + 1) mod 2N - 1) / 2) of
// NTT(x): slot p holds the evaluation at psi^(2 br(p) + 1).
template <int LOGN>
__device__ __forceinline__ uint32_t aut_slot(uint32_t p, uint32_t t) {
    constexpr uint32_t N = 1u << LOGN;
    const uint32_t e = (t * (2u * (__brev(p) >> (32 - LOGN)) + 1u)) & (2u * N - 1u);
    return __brev((e - 1u) >> 1) >> (32 - LOGN);
}

// Scales of the signed inverse transform of residue k (reg_ntt.hpp SInvScale):
// 2^32 mod q and the last stage's output scales in Montgomery form.
__device__ __forceinline__ SInvScale tree_sinv(const CtxDev& C, const ScaleDev& inv, int k) {
    const uint32_t q = C.m[k].q;
    SInvScale s;
    s.r1 = C.mont[k];
    s.cm = csub(shoup_mul(inv.c[k], C.mont[k], C.montp[k], q), q);
    s.cwm = csub(shoup_mul(inv.cw[k], C.mont[k], C.montp[k], q), q);
    return s;
}

// Logical block of hardware block b in a grid of n: the dispatcher hands
// consecutive blocks to the 8 XCDs in turn, each with its own L2, so the
// blocks XCD x receives (x, x + 8, ...) are given one contiguous range of
// logical ids.  The column workgroups of a query then share one L2 for the
// query's RGSW rows, which would otherwise be fetched once per XCD.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcd = 8;
    const uint32_t x = b % kXcd, i = b / kXcd, per = n / kXcd, rem = n % kXcd;
    return x * per + (x < rem ? x : rem) + i;
}

// One level of pack_lwes_inner on the node held by this workgroup (see the
// file comment).  a0/a1: s0.a/s1.a at this thread's coefficient positions,
// h0/h1: s0.a^/s1.a^ and b0/b1: s0.b/s1.b at its slot positions (canonical).
// On return a0, h0, b0 hold the new node (canonical).  key: [ell][2][K][N]
// NTT-domain Montgomery rows.
template <int LOGN, int K, int ELL, bool LAZY, bool REC, bool SGN>
__device__ __forceinline__ void tree_level(uint32_t (&a0)[kTreeV], const uint32_t (&a1)[kTreeV], uint32_t (&h0)[kTreeV],
                                           const uint32_t (&h1)[kTreeV], uint32_t (&b0)[kTreeV], const uint32_t (&b1)[kTreeV],
                                           uint32_t level, const uint32_t* key, const TreeLds<LOGN, K, ELL, REC>& L,
                                           const TreeArgs& TA, const CtxDev& C, const ScaleDev& inv, int k, int t) {
    using Ly = LayV<LOGN, kTreeV>;
    constexpr int V = kTreeV;
    constexpr int N = Ly::N, T = Ly::T, LP = Ly::LP;
    constexpr int ND = TreeDigits<ELL, REC>::ND, DB = TreeDigits<ELL, REC>::DB;
    constexpr size_t KN = (size_t)K * N;
    const Mod M{C.m[k].q, C.m[k].q2, C.m[k].qinv};
    const uint32_t q = M.q;
    const uint32_t taut = (1u << level) + 1u;
    const uint32_t tinv = inv_odd(taut) & (2u * N - 1u);

    // the key rows this residue multiplies by, loaded ahead of the transforms
    uint32_t ka[ELL][kTreeV], kb[ELL][kTreeV];
#pragma unroll
    for (int j = 0; j < ELL; j++) {
        load_slots_v<V>(key + ((size_t)j * 2 + 0) * KN + (size_t)k * N, t, ka[j]);
        load_slots_v<V>(key + ((size_t)j * 2 + 1) * KN + (size_t)k * N, t, kb[j]);
    }

    lds_barrier();  // every thread is past its reads of the previous level
    const int cix = lds_ix(t);
    uint32_t* ab = L.abuf + k * LP;
    uint32_t* bb = L.bbuf + k * LP;
    uint32_t* hb = L.hbuf + k * LP;
#pragma unroll
    for (int e = 0; e < V; e++) ab[lds_step(cix, e, T)] = a1[e];
#pragma unroll
    for (int g = 0; g < V / 4; g++) {
        *reinterpret_cast<uint4*>(&bb[lds_ix(Ly::group(t, g))]) =
            make_uint4(b1[4 * g + 0], b1[4 * g + 1], b1[4 * g + 2], b1[4 * g + 3]);
        if constexpr (REC)
            *reinterpret_cast<uint4*>(&hb[lds_ix(Ly::group(t, g))]) =
                make_uint4(h1[4 * g + 0], h1[4 * g + 1], h1[4 * g + 2], h1[4 * g + 3]);
    }
    lds_barrier();

    // gadget digits of tau_t(s1.a) at the coefficient positions:
    // tau_t(x)_j = x_i for t i = j mod 2N, and -x_i for t i = j + N mod 2N
    uint32_t dg[ELL][kTreeV];
#pragma unroll
    for (int e = 0; e < V; e++) {
        const uint32_t j = (uint32_t)(t + e * T);
        const uint32_t src = (tinv * j) & (2u * N - 1u);
        const bool neg = src >= (uint32_t)N;
        const int sx = lds_ix((int)(src & (N - 1)));
        uint32_t c[K];
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const uint32_t v = L.abuf[kk * LP + sx];
            c[kk] = neg ? negq(v, C.m[kk].q) : v;
        }
        uint64_t x = c[0];
        if constexpr (K == 2) {
            const uint32_t q1 = C.m[1].q;
            const uint32_t a0m = csub(shoup_mul(c[0], 1u, C.one_q1p, q1), q1);
            const uint32_t u = csub(shoup_mul(c[1] + q1 - a0m, C.q0_inv_q1, C.q0_inv_q1p, q1), q1);
            x = (uint64_t)c[0] + (uint64_t)C.m[0].q * u;
        }
#pragma unroll
        for (int j2 = 0; j2 < ND; j2++) {
            const uint32_t sh = j2 * C.gadget_log;
            const uint32_t v = sh < 64 ? (uint32_t)((x >> sh) & C.gadget_mask) : 0u;
            dg[j2][e] = C.small_digits ? v : reduce32(v, q, C.onep[k]);
        }
    }
    // tau_t(s1.b), and tau_t(s1.a^) for the recovered digit, at the slot positions
    uint32_t bt[kTreeV], ht[kTreeV];
#pragma unroll
    for (int g = 0; g < V / 4; g++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int sx = lds_ix((int)aut_slot<LOGN>((uint32_t)(Ly::group(t, g) + c), taut));
            bt[4 * g + c] = bb[sx];
            if constexpr (REC) ht[4 * g + c] = hb[sx];
        }

    if constexpr (ND > 0) {
        ntt_v<LOGN, V, ND, SGN>(reinterpret_cast<uint32_t(&)[ND][V]>(dg), t, L.dbuf + k * DB * LP, L.twf + k * N,
                                  M);
        // every thread is past its gathers from abuf before the INTT reuses it: the transform's
        // own barrier orders them when it has one
        if constexpr (!Ly::fwd_has_barrier) lds_barrier();
    } else {
        lds_barrier();
    }
    if constexpr (REC && SGN) {
        // NTT(D_last) = z^-(ell-1) tau(s1.a^) + sum_{j < ell-1} -z^(j-(ell-1)) NTT(D_j): one signed sum
#pragma unroll
        for (int e = 0; e < V; e++) {
            int64_t sum = (int64_t)(int32_t)ht[e] * (int32_t)TA.zinvm[k];
#pragma unroll
            for (int j = 0; j < ND; j++) sum += (int64_t)(int32_t)dg[j][e] * (int32_t)TA.zrecm[k][j];
            dg[ELL - 1][e] = smont(sum, M);
        }
    } else if constexpr (REC) {
        // transform outputs < 4q and Montgomery products < 2q keep the difference in (0, 7q)
#pragma unroll
        for (int e = 0; e < V; e++) {
            uint32_t d = ht[e];
            if constexpr (ND >= 1) d += 4 * M.q - dg[0][e];
#pragma unroll
            for (int j = 1; j < ND; j++) d += M.q2 - mmul(dg[j][e], TA.zpowm[k][j], M);
            dg[ELL - 1][e] = ND > 0 ? mmul(d, TA.zinvm[k], M) : d;
        }
    }

    uint32_t A[kTreeV];
#pragma unroll
    for (int e = 0; e < V; e++) {
        if constexpr (SGN) {
            int64_t sa = 0, sb = 0;
#pragma unroll
            for (int j = 0; j < ELL; j++) {
                const int64_t x = (int32_t)dg[j][e];
                sa += x * (int32_t)ka[j][e];
                sb += x * (int32_t)kb[j][e];
            }
            // reduced sums in (-2q, 2q): + 2q and one conditional subtraction give [0, 2q)
            A[e] = csub(smont(sa, M) + M.q2, M.q2);
            h0[e] = addq(h0[e], csub(A[e], q), q);
            b0[e] = addq(addq(b0[e], bt[e], q), csub(csub(smont(sb, M) + M.q2, M.q2), q), q);
        } else {
            uint64_t sa = 0, sb = 0;
#pragma unroll
            for (int j = 0; j < ELL; j++) {
                const uint32_t x = LAZY ? dg[j][e] : csub(csub(dg[j][e], M.q2), q);
                sa += (uint64_t)x * ka[j][e];
                sb += (uint64_t)x * kb[j][e];
            }
            A[e] = mont_reduce(sa, q, M.qinv);  // < 2q, the inverse transform's input range
            h0[e] = addq(h0[e], csub(A[e], q), q);
            b0[e] = addq(addq(b0[e], bt[e], q), csub(mont_reduce(sb, q, M.qinv), q), q);
        }
    }
    if constexpr (SGN) {
        // signed inverse transform (A in [0, 2q), 8 q < 2^31 checked on the host): one
        // reduction per radix-4 pass, output in (-q, q)
        intt4_s<LOGN>(A, t, ab, L.twf + k * N, M, tree_sinv(C, inv, k));
#pragma unroll
        for (int e = 0; e < V; e++) a0[e] = addq(a0[e], sfix(A[e], q), q);
    } else {
        intt_v<LOGN, V>(A, t, ab, L.twf + k * N, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
#pragma unroll
        for (int e = 0; e < V; e++) a0[e] = addq(a0[e], csub(A[e], q), q);
    }
}

template <int LOGN, int K, int ELL, bool LAZY, bool REC, bool SGN>
__global__ void __launch_bounds__(K * ((1 << LOGN) / kTreeV), 1)
k_tree_leaves(TreeArgs A, CtxDev C, ScaleDev inv) {
    using Ly = LayV<LOGN, kTreeV>;
    constexpr int V = kTreeV;
    constexpr int N = Ly::N, T = Ly::T, LP = Ly::LP;
    constexpr size_t KN = (size_t)K * N;
    extern __shared__ __align__(16) uint32_t lds[];
    const TreeLds<LOGN, K, ELL, REC> L(lds);
    const int tid = threadIdx.x, k = residue_of<T>(tid), t = tid - k * T;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t qi = blk / A.m, s = blk - qi * A.m;
    const Mod M{C.m[k].q, C.m[k].q2, C.m[k].qinv};
    const uint32_t q = M.q;

    // first dimension of column s (external_product of (0, Delta h) with RGSW
    // rows ell..2ell-1; the a-digits are 0): NTT domain, slot positions
    const uint32_t* dig = reinterpret_cast<const uint32_t*>(A.qdig[qi]) + (size_t)s * ELL * KN + (size_t)k * N;
    const uint32_t* rw = A.rows + (size_t)qi * ELL * 2 * KN + (size_t)k * N;
    uint32_t dv[ELL][kTreeV], ra[ELL][kTreeV], rb[ELL][kTreeV];
#pragma unroll
    for (int j = 0; j < ELL; j++) {
        load_slots_v_nt<V>(dig + (size_t)j * KN, t, dv[j]);
        load_slots_v<V>(rw + (size_t)j * 2 * KN, t, ra[j]);
        load_slots_v<V>(rw + (size_t)j * 2 * KN + KN, t, rb[j]);
    }
    stage_twiddles(L, C);
    uint32_t ah[kTreeV];
    uint64_t bsum = 0;
#pragma unroll
    for (int e = 0; e < V; e++) {
        uint64_t sa = 0, sb = 0;
#pragma unroll
        for (int j = 0; j < ELL; j++) {
            sa += (uint64_t)dv[j][e] * ra[j][e];
            sb += (uint64_t)dv[j][e] * rb[j][e];
        }
        ah[e] = mont_reduce(sa, q, M.qinv);
        bsum += csub(mont_reduce(sb, q, M.qinv), q);
    }
    // LWE b = coefficient 0 of the column's b = n^-1 * sum of its slots
    constexpr int W = TreeShape<LOGN, K, ELL, REC>::W;  // a group never spans two residues
#pragma unroll
    for (int o = W / 2; o >= 1; o >>= 1) bsum += __shfl_xor(bsum, o, W);
    if (tid % W == 0) L.red[tid / W] = bsum;
    __syncthreads();  // also orders the twiddle stores before their first use
    uint64_t tot = 0;
#pragma unroll
    for (int w = 0; w < T / W; w++) tot += L.red[k * (T / W) + w];
    // (tot 2^-32) (n^-1 2^32) = tot n^-1; tot < N q < q 2^32
    const uint32_t lwe_b =
        csub(shoup_mul(csub(mont_reduce(tot, q, M.qinv), q), A.ninvm[k], A.ninvmp[k], q), q);

    // the leaf: (a, 0) with a the column's a (invert_sample_extract undoes
    // sample_extract_coeff0, automorph_pack.rs:349-405)
    uint32_t a[kTreeV], h[kTreeV], b[kTreeV];
#pragma unroll
    for (int e = 0; e < V; e++) {
        h[e] = csub(ah[e], q);
        a[e] = h[e];
        b[e] = 0;
    }
    if constexpr (SGN) {
        intt4_s<LOGN>(a, t, L.abuf + k * LP, L.twf + k * N, M, tree_sinv(C, inv, k));  // a < 2q in
#pragma unroll
        for (int e = 0; e < V; e++) a[e] = sfix(a[e], q);
    } else {
        intt_v<LOGN, V>(a, t, L.abuf + k * LP, L.twf + k * N, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
#pragma unroll
        for (int e = 0; e < V; e++) a[e] = csub(a[e], q);
    }

    if (A.m == 1) {  // a single LWE is not tree-packed (automorph_pack.rs:436-443)
        uint32_t* o = A.out + (size_t)qi * A.out_words;
#pragma unroll
        for (int e = 0; e < V; e++) {
            const int j = t + e * T;
            o[(size_t)k * N + j] = a[e];
            o[KN + (size_t)k * N + j] = j == 0 ? lwe_b : 0u;
        }
        return;
    }
    if (t == 0) A.lweb[((size_t)qi * A.m + s) * K + k] = lwe_b;

    for (uint32_t lv = A.level_lo; lv <= A.level_hi; lv++)
        tree_level<LOGN, K, ELL, LAZY, REC, SGN>(a, a, h, h, b, b, lv, A.ks + (size_t)(A.logn - lv) * ELL * 2 * KN, L, A,
                                            C, inv, k, t);

    uint32_t* nd = A.node + ((size_t)qi * A.m + s) * 3 * KN + (size_t)k * N;
#pragma unroll
    for (int e = 0; e < V; e++) nd[t + e * T] = a[e];
#pragma unroll
    for (int g = 0; g < V / 4; g++) {
        *reinterpret_cast<uint4*>(nd + KN + Ly::group(t, g)) = make_uint4(h[4 * g], h[4 * g + 1], h[4 * g + 2], h[4 * g + 3]);
        *reinterpret_cast<uint4*>(nd + 2 * KN + Ly::group(t, g)) =
            make_uint4(b[4 * g], b[4 * g + 1], b[4 * g + 2], b[4 * g + 3]);
    }
}

// One level after the leaves: node s of level `level` from nodes s and
// s + 2^(log d - level) of the level below (in place: outputs s <
// 2^(log d - level) never hold an odd input).  At the root the packed
// ciphertext is finished into out.
template <int LOGN, int K, int ELL, bool LAZY, bool REC, bool SGN>
__global__ void __launch_bounds__(K * ((1 << LOGN) / kTreeV), 1)
k_tree_combine(TreeArgs A, CtxDev C, ScaleDev inv) {
    using Ly = LayV<LOGN, kTreeV>;
    constexpr int V = kTreeV;
    constexpr int N = Ly::N, T = Ly::T, LP = Ly::LP;
    constexpr size_t KN = (size_t)K * N;
    extern __shared__ __align__(16) uint32_t lds[];
    const TreeLds<LOGN, K, ELL, REC> L(lds);
    const int tid = threadIdx.x, k = residue_of<T>(tid), t = tid - k * T;
    const uint32_t level = A.level_lo;
    const uint32_t half = 1u << (LOGN - level);  // nodes at this level: s < min(m, half)
    const uint32_t nn = A.m < half ? A.m : half;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t qi = blk / nn, s = blk - qi * nn;
    const bool odd = s + half < A.m;
    const uint32_t q = C.m[k].q;
    const uint32_t* ev = A.node + ((size_t)qi * A.m + s) * 3 * KN + (size_t)k * N;
    const uint32_t* od = A.node + ((size_t)qi * A.m + s + half) * 3 * KN + (size_t)k * N;

    uint32_t a0[kTreeV], a1[kTreeV], h0[kTreeV], h1[kTreeV], b0[kTreeV], b1[kTreeV];
    const uint32_t ysh = (uint32_t)N >> level;  // y = X^(N / 2^level)
#pragma unroll
    for (int e = 0; e < V; e++) {
        const uint32_t j = (uint32_t)(t + e * T);
        const uint32_t x = ev[j];
        uint32_t v = 0;
        if (odd) v = j >= ysh ? od[j - ysh] : negq(od[j + N - ysh], q);  // (X^ysh odd.a)_j
        a0[e] = addq(x, v, q);
        a1[e] = subq(x, v, q);
    }
    {
        // a^ and b: s0 = even + y^ odd, s1 = even - y^ odd, slot by slot
        uint32_t eh[kTreeV], oh[kTreeV], eb[kTreeV], ob[kTreeV], yh[kTreeV];
        load_slots_v_nt<V>(ev + KN, t, eh);
        load_slots_v_nt<V>(ev + 2 * KN, t, eb);
        if (odd) {
            load_slots_v_nt<V>(od + KN, t, oh);
            load_slots_v_nt<V>(od + 2 * KN, t, ob);
            load_slots_v<V>(A.yhat + ((size_t)(level - 1) * K + k) * N, t, yh);
        }
        const Mod M{q, C.m[k].q2, C.m[k].qinv};
#pragma unroll
        for (int e = 0; e < V; e++) {
            const uint32_t vh = odd ? csub(mmul(oh[e], yh[e], M), q) : 0u;
            h0[e] = addq(eh[e], vh, q);
            h1[e] = subq(eh[e], vh, q);
            const uint32_t v = odd ? csub(mmul(ob[e], yh[e], M), q) : 0u;
            b0[e] = addq(eb[e], v, q);
            b1[e] = subq(eb[e], v, q);
        }
    }
    stage_twiddles(L, C);  // the level's first barrier orders these stores
    tree_level<LOGN, K, ELL, LAZY, REC, SGN>(a0, a1, h0, h1, b0, b1, level, A.ks + (size_t)(A.logn - level) * ELL * 2 * KN,
                                        L, A, C, inv, k, t);

    if (level < (uint32_t)LOGN) {
        // each thread rewrites exactly the words of node s it read
        uint32_t* nd = A.node + ((size_t)qi * A.m + s) * 3 * KN + (size_t)k * N;
#pragma unroll
        for (int e = 0; e < V; e++) nd[t + e * T] = a0[e];
#pragma unroll
        for (int g = 0; g < V / 4; g++) {
            *reinterpret_cast<uint4*>(nd + KN + Ly::group(t, g)) =
                make_uint4(h0[4 * g], h0[4 * g + 1], h0[4 * g + 2], h0[4 * g + 3]);
            *reinterpret_cast<uint4*>(nd + 2 * KN + Ly::group(t, g)) =
                make_uint4(b0[4 * g], b0[4 * g + 1], b0[4 * g + 2], b0[4 * g + 3]);
        }
        return;
    }

    // root: b to the coefficient domain, then b_val * d at coefficient z < m
    const Mod M{q, C.m[k].q2, C.m[k].qinv};
    lds_barrier();  // every thread is past the level's last reads of abuf
    intt_v<LOGN, V>(b0, t, L.abuf + k * LP, L.twf + k * N, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
    uint32_t* o = A.out + (size_t)qi * A.out_words;
    uint32_t* ob = o + KN;
    const uint32_t* lw = A.lweb + (size_t)qi * A.m * K;
#pragma unroll
    for (int e = 0; e < V; e++) {
        const uint32_t j = (uint32_t)(t + e * T);
        o[(size_t)k * N + j] = a0[e];
        const uint32_t bv = csub(b0[e], q);
        if (!A.ref_layout) {
            uint32_t v = bv;
            if (j < A.m) v = addq(v, csub(shoup_mul(lw[(size_t)j * K + k], A.dmod[k], A.dmodp[k], q), q), q);
            ob[(size_t)k * N + j] = v;
        } else if (k == 0) {
            // b.coeffs() = [res0 | res1] read as 2N composite coefficients
            // (Poly::from_coeffs_moduli, poly.rs:127-159): word z < N is
            // (b0_z + b_val_z d) mod q, cut into both residues
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                const uint32_t qk = C.m[kk].q;
                uint32_t v = kk == 0 ? bv : reduce32(bv, qk, C.onep[kk]);
                if (j < A.m) v = addq(v, csub(shoup_mul(lw[(size_t)j * K + kk], A.dmod[kk], A.dmodp[kk], qk), qk), qk);
                ob[(size_t)kk * 2 * N + j] = v;
            }
        } else {
            // word N + j is b1_j, below q1, cut into both residues
#pragma unroll
            for (int kk = 0; kk < K; kk++)
                ob[(size_t)kk * 2 * N + N + j] = kk == 1 ? bv : reduce32(bv, C.m[kk].q, C.onep[kk]);
        }
    }
}

template <int LOGN, int K, int ELL, bool LAZY, bool REC, bool SGN = false>
static hipError_t tree_launch(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                              hipStream_t st) {
    constexpr size_t shmem = TreeShape<LOGN, K, ELL, REC>::kBytes;
    static_assert(shmem <= 160 * 1024, "tree kernel LDS");
    auto kern = leaves ? k_tree_leaves<LOGN, K, ELL, LAZY, REC, SGN> : k_tree_combine<LOGN, K, ELL, LAZY, REC, SGN>;
    // raise the dynamic LDS limit once per instantiation
    static const hipError_t e0 =
        shmem > 64 * 1024 ? hipFuncSetAttribute((const void*)k_tree_leaves<LOGN, K, ELL, LAZY, REC, SGN>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)
                          : hipSuccess;
    static const hipError_t e1 =
        shmem > 64 * 1024 ? hipFuncSetAttribute((const void*)k_tree_combine<LOGN, K, ELL, LAZY, REC, SGN>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)
                          : hipSuccess;
    if (e0 != hipSuccess) return e0;
    if (e1 != hipSuccess) return e1;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(K * ((1 << LOGN) / kTreeV)), shmem, st, A, C, inv);
    return hipGetLastError();
}

template <int LOGN, int K, int ELL>
static hipError_t tree_lazy(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                            hipStream_t st) {
    uint64_t qmax = 0;
    for (int kk = 0; kk < K; kk++) qmax = C.m[kk].q > qmax ? C.m[kk].q : qmax;
    // the recovered digit needs the composite value to fit ell digits: q <= z^ell
    const bool rec = (uint64_t)C.gadget_log * ELL >= 64 || C.q <= (1ull << (C.gadget_log * ELL));
    // signed digit path (ct_s transforms, signed Montgomery sums): tau(s1.a^) < q enters the
    // recovered digit, nothing is folded into the multiply-accumulates
    // the signed inverse transforms hold sums below 8 q in int32
    bool sgn = 8 * qmax < (1ull << 31);
    for (int kk = 0; kk < K; kk++) {
        const uint64_t q = C.m[kk].q;
        sgn = sgn && signed_digits_ok(q, C.small_digits ? C.gadget_mask : q - 1, LOGN, rec ? ELL - 1 : ELL,
                                      rec ? q : 0, 0);
    }
    if ((uint64_t)ELL * 4 * qmax < (1ull << 32)) {
        if (sgn && rec) return tree_launch<LOGN, K, ELL, true, true, true>(A, leaves, blocks, C, inv, st);
        if (sgn) return tree_launch<LOGN, K, ELL, true, false, true>(A, leaves, blocks, C, inv, st);
        if (rec) return tree_launch<LOGN, K, ELL, true, true>(A, leaves, blocks, C, inv, st);
        return tree_launch<LOGN, K, ELL, true, false>(A, leaves, blocks, C, inv, st);
    }
    if (rec) return tree_launch<LOGN, K, ELL, false, true>(A, leaves, blocks, C, inv, st);
    return tree_launch<LOGN, K, ELL, false, false>(A, leaves, blocks, C, inv, st);
}

template <int LOGN, int K>
static hipError_t tree_ell(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                           hipStream_t st) {
    switch (C.ell) {
        case 1: return tree_lazy<LOGN, K, 1>(A, leaves, blocks, C, inv, st);
        case 2: return tree_lazy<LOGN, K, 2>(A, leaves, blocks, C, inv, st);
        case 3: return tree_lazy<LOGN, K, 3>(A, leaves, blocks, C, inv, st);
        default: return hipErrorInvalidValue;
    }
}

template <int LOGN>
static hipError_t tree_k(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                         hipStream_t st) {
    if (C.K == 1) return tree_ell<LOGN, 1>(A, leaves, blocks, C, inv, st);
    if (C.K == 2) return tree_ell<LOGN, 2>(A, leaves, blocks, C, inv, st);
    return hipErrorInvalidValue;
}

}  // namespace

bool tree_supported(const CtxDev& C) {
    if (C.K < 1 || C.K > 2 || C.ell < 1 || C.ell > 3 || C.logn < 8 || C.logn > 11) return false;
    for (uint32_t k = 0; k < C.K; k++)
        if (C.m[k].q >= (1u << 30)) return false;  // lazy ranges: 4q < 2^32
    return true;
}

hipError_t launch_tree(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                       hipStream_t st) {
    if (!blocks) return hipSuccess;
    switch (C.logn) {
        case 8: return tree_k<8>(A, leaves, blocks, C, inv, st);
        case 9: return tree_k<9>(A, leaves, blocks, C, inv, st);
        case 10: return tree_k<10>(A, leaves, blocks, C, inv, st);
        case 11: return tree_k<11>(A, leaves, blocks, C, inv, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
