// Packing chain for two CRT moduli: packing_offline step 2 and the online sum
// (src/inspiring/inspiring2.rs:926-1041, :1185-1222), one query per
// workgroup of 2 x N/V threads (V values each, kChainV below); the first half
// works on residue 0, the second on residue 1.
//
// Per step i = gamma-2 .. 0, in the NTT domain:
//   T[i] = G^-1(INTT(R[i+1]));  R[i] = tau_{g^i}(r_pre[i]) + sum_j NTT(T[i][j]) * w_all[i][j]
//   S   += sum_j NTT(T[i][j]) * y_all[i][j]          (online accumulator)
// then a_hat = INTT(R[0]) and b = b_poly + INTT(S).
//
// Thread layout.  The register-resident NTT passes (radix 4 or 8) leave every
// thread with V fixed positions.  The inverse transform's last pass and the
// forward transform's first pass touch the same positions {t + k N/V} (the
// "coefficient positions"), and the forward transform's last pass and the
// inverse transform's first pass touch the same positions (the "slot
// positions", V t .. V t + V - 1).  So, per step, the CRT composition and the
// gadget digits of T[i] are formed in registers between the two transforms
// (after one LDS exchange of the coefficients between the halves), and the
// multiply-accumulate that produces R[i] runs in registers on the slots the
// next inverse transform starts from.  R, the online accumulator S and the
// NTT of the client's y_body (rotated by the automorphism table on the fly,
// staged in LDS) never leave the chip.
//
// Twiddles are kept in LDS in Montgomery form (w 2^32 mod q, one word each),
// the forward table only (the inverse transforms read it mirrored,
// reg_ntt.hpp); a butterfly's product is one Montgomery reduction.  Barriers wait for LDS
// traffic only, so the global loads of the next multiply-accumulate's
// operands (w_all, the r_pre gathers, the permutation) stay in flight across
// the transforms.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"
#include "ntt.hpp"
#include "pack.hpp"
#include "reg_ntt.hpp"

namespace inspire {

// REC: the last gadget digit's transform is recovered instead of computed.
// The digits D_j of T = INTT(R) satisfy T = sum_j z^j D_j exactly when
// z^ell >= q (the last digit is never truncated), and the NTT is linear mod
// q_k, so NTT(D_last) = z^-(ell-1) (R - sum_{j<ell-1} z^j NTT(D_j)) mod q_k:
// two modular products per slot replace a transform, and one LDS digit
// buffer is not needed.
template <int ELL, bool REC>
struct ChainDigits {
    static constexpr int ND = REC ? ELL - 1 : ELL;  // digits transformed
    static constexpr int DB = ND > 0 ? ND : 1;      // LDS exchange buffers per residue
};

template <int LOGN, int ELL, bool REC, int V>
struct ChainCrtShape {
    static constexpr int N = 1 << LOGN, LP = LayV<LOGN, V>::LP;
    static constexpr size_t kBytes = (size_t)2 * N * 4      // forward Montgomery twiddles of both residues
                                     + (size_t)ELL * 2 * N * 4  // NTT of the client's y_body
                                     + (size_t)2 * ChainDigits<ELL, REC>::DB * LP * 4  // transform exchange,
                                     // per residue (the inverse uses the first digit buffer)
                                     + (size_t)2 * LP * 4;  // coefficient exchange between the halves
};

// Position of NTT slot s of the client's keys in LDS.  The rotation
// tau_{g^i} with g = 2n/gamma + 1 moves a slot only within its aligned group
// of 16 (it changes the top bits of the slot's bit-reversed index), by the
// same map for every group of a wave's 256 slots, so in the plain layout the
// gathers of a wave's groups all hit the same 4 banks of their parity.  The
// XOR of bits 0-1 with bits 5-6 and 7-8 gives each such group its own banks
// (32 banks per half wave or 64 per wave, checked by simulating every
// rotation's gathers); it is a bijection, so any permutation stays correct.
__device__ __forceinline__ int yl_ix(int s) { return s ^ (((s >> 5) ^ (s >> 7)) & 3); }

// LAZY: ELL * 4q * q < q * 2^32, so the multiply-accumulate takes the
// transform outputs (< 4q) without reducing them first.
// SGN: the digit transforms run on two's-complement words with butterflies
// that reduce nothing (ct_s), the recovered digit and the multiply-accumulate
// (into which R's and S's previous values are folded) are signed 64-bit sums
// with one signed Montgomery reduction each, and R, S stay in [0, 2q).  The
// host enables it only when chain_signed_ok() shows that every transform
// stage stays inside int32, every reduced sum inside (-2q, 2q), and q0 < 2 q1.
template <int LOGN, int ELL, bool LAZY, bool REC, bool SGN, int V>
__global__ void __launch_bounds__(2 * ((1 << LOGN) / V), 1)
k_pack_chain_crt(PackArgs A, CtxDev C, ScaleDev inv) {
    using L = LayV<LOGN, V>;
    constexpr int N = L::N, T = N / V, LP = L::LP;  // T threads per residue, V values each
    constexpr int ND = ChainDigits<ELL, REC>::ND, DB = ChainDigits<ELL, REC>::DB;
    constexpr size_t KN = 2 * (size_t)N;
    extern __shared__ __align__(16) uint32_t lds[];
    uint32_t* twf = lds;                 // [2][N] psi^br(i) 2^32 mod q_k
    uint32_t* yl = twf + 2 * N;          // [ELL][2][N]
    uint32_t* dbuf = yl + ELL * 2 * N;   // [2][DB][LP]
    uint32_t* xbuf = dbuf + 2 * DB * LP;  // [2][LP]

    const int tid = threadIdx.x;
    const int k = residue_of<T>(tid), t = tid - k * T;  // residue of this half, thread within it
    const uint32_t qi = blockIdx.x, gamma = A.gamma;
    const uint32_t* rpre = A.r_pre + (size_t)qi * gamma * KN;
    const uint32_t* Y = A.y + (A.y_off ? A.y_off[qi] : (uint64_t)qi * ELL * KN);
    const uint32_t ymode = A.y_mode ? A.y_mode[qi] : 0u;
    const uint32_t ycount = A.y_count ? A.y_count[qi] : gamma - 1;
    const Mod M{C.m[k].q, C.m[k].q2, C.m[k].qinv};
    const uint32_t q0 = C.m[0].q, q1 = C.m[1].q;
    // SINV: the steps' inverse transforms are signed (intt4_s; the host's
    // chain_signed_ok also requires 8 q < 2^31)
    constexpr bool SINV = SGN && V == 4;
    SInvScale sinv{};
    if constexpr (SINV) {
        sinv.r1 = C.mont[k];
        sinv.cm = csub(shoup_mul(inv.c[k], C.mont[k], C.montp[k], M.q), M.q);
        sinv.cwm = csub(shoup_mul(inv.cw[k], C.mont[k], C.montp[k], M.q), M.q);
    }
    uint32_t* db = dbuf + k * DB * LP;
    uint32_t* rb = db;  // the inverse transform's exchange: this residue's first digit buffer
    const uint32_t* tw = twf + k * N;

    for (int e = tid; e < 2 * N; e += 2 * T) {
        const int kk = e / N, i = e - kk * N;
        const ModDev& D = C.m[kk];
        twf[e] = csub(shoup_mul(D.tw[i], C.mont[kk], C.montp[kk], D.q), D.q);
    }
    if (A.roles.role && A.roles.role[qi]) {
        // InsPIRe^0 in a packing launch: the front end left the query's
        // combined and column ciphertexts [1 + gamma][2][K][N] in the NTT
        // domain; each half transforms its residue of every poly in place
        // (the next poly's slots loaded while the current one transforms)
        __syncthreads();  // the twiddles
        uint32_t* O = A.out + A.roles.qout[qi] + (size_t)k * N;
        const uint32_t np = 2 * (1 + gamma);
        uint32_t v[V], nx[V];
        load_slots_v<V>(O, t, v);
        for (uint32_t p = 0; p < np; p++) {
            if (p + 1 < np) load_slots_v<V>(O + (size_t)(p + 1) * KN, t, nx);
            if (p) lds_barrier();  // every thread is past the previous transform's reads of rb
            intt_v<LOGN, V>(v, t, rb, tw, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
            // the transform's barriers order every thread's slot loads of poly p before these stores
#pragma unroll
            for (int e = 0; e < V; e++) {
                O[(size_t)p * KN + t + e * T] = csub(v[e], M.q);
                v[e] = nx[e];
            }
        }
        return;
    }
    if (ymode == 0 && A.y_coeff) {
        // NTT of this query's y_body (both halves, one residue each), into
        // LDS at the slot positions in Montgomery form
        __syncthreads();
        // DB polys at a time through the DB exchange buffers, the last group
        // only as many as are left
        auto y_group = [&](auto cnt, int j0) {
            constexpr int NP = decltype(cnt)::value;
            uint32_t yv[NP][V];
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int e = 0; e < V; e++) yv[j][e] = Y[((j0 + j) * 2 + k) * N + t + e * T];
            if (j0 > 0) lds_barrier();  // every thread is past the previous transform's reads
            ntt_v<LOGN, V, NP>(yv, t, db, tw, M);
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int e = 0; e < V; e++)
                    yl[((j0 + j) * 2 + k) * N + yl_ix(V * t + e)] =
                        csub(shoup_mul(yv[j][e], C.mont[k], C.montp[k], M.q), M.q);
        };
#pragma unroll
        for (int j0 = 0; j0 + DB <= ELL; j0 += DB) y_group(std::integral_constant<int, DB>{}, j0);
        if constexpr (ELL % DB != 0) y_group(std::integral_constant<int, ELL % DB>{}, ELL - ELL % DB);
    } else if (ymode == 0) {
        for (int e = tid; e < ELL * 2 * N; e += 2 * T) yl[(e & ~(N - 1)) + yl_ix(e & (N - 1))] = Y[e];
    }

    // this residue's value of the reference's automorphism of r_pre[i] at a
    // slot whose source slot is src (inspiring2.rs:279-296): residue 1 is
    // rebuilt from residue 0's value in the reference's 2^64 Montgomery form,
    // or permuted on its own in the corrected mode
    auto r_init = [&](const uint32_t* Ri, uint32_t src) -> uint32_t {
        const uint32_t v0 = Ri[src];
        if (k == 0) return v0;
        if (A.residue_wise) return Ri[N + src];
        const uint32_t ref = csub(shoup_mul(v0, A.q0r64, A.q0r64p, q0), q0);
        const uint32_t v1 = csub(shoup_mul(ref, 1u, A.one_q1p, q1), q1);
        return csub(shoup_mul(v1, A.q1r64i, A.q1r64ip, q1), q1);
    };

    uint32_t R[V], S[V], pm[V], pmn[V];
    load_slots_v<V>(A.perm + (size_t)(gamma - 1) * N, t, pm);
#pragma unroll
    for (int e = 0; e < V; e++) {
        R[e] = r_init(rpre + (size_t)(gamma - 1) * KN, pm[e]);
        S[e] = 0;
    }
    if (gamma >= 2) load_slots_v<V>(A.perm + (size_t)(gamma - 2) * N, t, pmn);
    __syncthreads();

    for (int i = (int)gamma - 2; i >= 0; i--) {
        // The transforms' LDS and twiddle addresses depend on t alone; the
        // compiler keeps them in registers across the steps (112 VGPRs, still
        // four waves per SIMD): 675 instead of 752 vector instructions a step.
        const int ts = t;
#pragma unroll
        for (int e = 0; e < V; e++) pm[e] = pmn[e];
        if (i > 0) load_slots_v<V>(A.perm + (size_t)(i - 1) * N, ts, pmn);
        // operands of this step's multiply-accumulate, loaded ahead
        const uint32_t* Ri = rpre + (size_t)i * KN;
        uint32_t rv[V], w[ELL][V];
#pragma unroll
        for (int e = 0; e < V; e++) rv[e] = Ri[pm[e]];
        uint32_t rv1[V];
        if (k == 1 && A.residue_wise) {
#pragma unroll
            for (int e = 0; e < V; e++) rv1[e] = Ri[N + pm[e]];
        }
        const uint32_t* Wi = A.w_all + (size_t)i * ELL * KN + (size_t)k * N;
#pragma unroll
        for (int j = 0; j < ELL; j++) load_slots_v<V>(Wi + (size_t)j * KN, ts, w[j]);

        // T[i] = G^-1(INTT(R[i+1])): this residue's coefficients, then the other's via LDS
        uint32_t Rs[V];  // R[i+1] itself, for the recovered digit
        if constexpr (REC) {
#pragma unroll
            for (int e = 0; e < V; e++) Rs[e] = R[e];
        }
        if constexpr (SINV) {
            intt4_s<LOGN>(R, ts, rb, tw, M, sinv);
#pragma unroll
            for (int e = 0; e < V; e++) R[e] = sfix(R[e], M.q);
        } else {
            intt_v<LOGN, V>(R, ts, rb, tw, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
#pragma unroll
            for (int e = 0; e < V; e++) R[e] = csub(R[e], M.q);
        }
        const int cix = lds_ix(ts);
        // the halves swap coefficients through buffers of their own, so the digit
        // transforms may write their exchange buffers while the other half still reads
        uint32_t* xb = xbuf + k * LP;
#pragma unroll
        for (int e = 0; e < V; e++) xb[lds_step(cix, e, T)] = R[e];
        lds_barrier();
        const uint32_t* ro = xbuf + (1 - k) * LP;  // the other residue's coefficients
        uint32_t dg[ELL][V];
#pragma unroll
        for (int e = 0; e < V; e++) {
            const uint32_t other = ro[lds_step(cix, e, T)];
            const uint32_t c0 = k == 0 ? R[e] : other, c1 = k == 0 ? other : R[e];
            // c0 mod q1 (SGN also guarantees q0 < 2 q1)
            const uint32_t a0m = SGN ? csub(c0, q1) : csub(shoup_mul(c0, 1u, C.one_q1p, q1), q1);
            const uint32_t u = csub(shoup_mul(c1 + q1 - a0m, C.q0_inv_q1, C.q0_inv_q1p, q1), q1);
            const uint64_t x = (uint64_t)c0 + (uint64_t)q0 * u;
#pragma unroll
            for (int j = 0; j < ND; j++) {
                const uint32_t sh = j * C.gadget_log;
                const uint32_t v = sh < 64 ? (uint32_t)((x >> sh) & C.gadget_mask) : 0u;
                dg[j][e] = (SGN || C.small_digits) ? v : reduce32(v, M.q, C.onep[k]);  // SGN implies small digits
            }
        }
        if constexpr (ND > 0) ntt_v<LOGN, V, ND, SGN>(reinterpret_cast<uint32_t(&)[ND][V]>(dg), ts, db, tw, M);
        if constexpr (REC && SGN) {
            // NTT(D_last) = z^-(ell-1) R + sum_{j < ell-1} -z^(j-(ell-1)) NTT(D_j): one signed sum
#pragma unroll
            for (int e = 0; e < V; e++) {
                int64_t sum = (int64_t)(int32_t)Rs[e] * (int32_t)A.zinvm[k];
#pragma unroll
                for (int j = 0; j < ND; j++) sum += (int64_t)(int32_t)dg[j][e] * (int32_t)A.zrecm[k][j];
                dg[ELL - 1][e] = smont(sum, M);
            }
        } else if constexpr (REC) {
            // NTT(D_last) = z^-(ell-1) (R - sum_{j < ell-1} z^j NTT(D_j)): transform outputs are
            // below 4q, the Montgomery products below 2q, so the difference stays in (0, 7q)
#pragma unroll
            for (int e = 0; e < V; e++) {
                uint32_t d = Rs[e];
                if constexpr (ND >= 1) d += 4 * M.q - dg[0][e];
#pragma unroll
                for (int j = 1; j < ND; j++) d += M.q2 - mmul(dg[j][e], A.zpowm[k][j], M);
                dg[ELL - 1][e] = ND > 0 ? mmul(d, A.zinvm[k], M) : d;
            }
        }

        const bool use_y = (uint32_t)i < ycount;
        uint32_t yv[ELL][V];
        if (use_y) {
            if (ymode == 0) {
                int ym[V];  // the gathered slots' positions in the swizzled layout
#pragma unroll
                for (int e = 0; e < V; e++) ym[e] = yl_ix((int)pm[e]);
#pragma unroll
                for (int j = 0; j < ELL; j++)
#pragma unroll
                    for (int e = 0; e < V; e++) yv[j][e] = yl[(j * 2 + k) * N + ym[e]];
            } else {
#pragma unroll
                for (int j = 0; j < ELL; j++)
                    load_slots_v<V>(Y + (((size_t)i * ELL + j) * 2 + k) * N, ts, yv[j]);
            }
        }
        // this residue's r_pre values: one wave-uniform choice per step, not per value
        uint32_t ri[V];
        if (k == 0) {
#pragma unroll
            for (int e = 0; e < V; e++) ri[e] = rv[e];
        } else if (A.residue_wise) {
#pragma unroll
            for (int e = 0; e < V; e++) ri[e] = rv1[e];
        } else {
#pragma unroll
            for (int e = 0; e < V; e++) {
                // r_pre's value enters the sum as r * rmul: residue 1 of the reference's automorphism
                // is residue 0's 2^64 Montgomery value mod q1, times 2^-64 (rmul = 2^-32 mod q1)
                if constexpr (SGN) {
                    ri[e] = csub(csub(shoup_mul(rv[e], A.q0r64, A.q0r64p, q0), q0), q1);
                } else {
                    const uint32_t ref = csub(shoup_mul(rv[e], A.q0r64, A.q0r64p, q0), q0);
                    const uint32_t v1 = csub(shoup_mul(ref, 1u, A.one_q1p, q1), q1);
                    ri[e] = csub(shoup_mul(v1, A.q1r64i, A.q1r64ip, q1), q1);
                }
            }
        }
        // the multiply-accumulates, with and without the online sum (use_y is
        // wave-uniform; one branch per step)
        auto mac = [&](auto with_y) {
            constexpr bool WY = decltype(with_y)::value;
#pragma unroll
            for (int e = 0; e < V; e++) {
                if constexpr (SGN) {
                    int64_t ar = (int64_t)(int32_t)ri[e] * (int32_t)A.rmul[k];
                    int64_t ab = (int64_t)(int32_t)S[e] * (int32_t)A.smul[k];
#pragma unroll
                    for (int j = 0; j < ELL; j++) {
                        const int64_t x = (int32_t)dg[j][e];
                        ar += x * (int32_t)w[j][e];
                        if constexpr (WY) ab += x * (int32_t)yv[j][e];
                    }
                    // reduced sums in (-2q, 2q): + 2q, one conditional subtraction -> [0, 2q)
                    R[e] = csub(smont(ar, M) + M.q2, M.q2);
                    if constexpr (WY) S[e] = csub(smont(ab, M) + M.q2, M.q2);
                } else {
                    uint64_t ar = 0, ab = 0;
#pragma unroll
                    for (int j = 0; j < ELL; j++) {
                        const uint32_t x = LAZY ? dg[j][e] : csub(csub(dg[j][e], M.q2), M.q);
                        ar += (uint64_t)x * w[j][e];
                        if constexpr (WY) ab += (uint64_t)x * yv[j][e];
                    }
                    R[e] = csub(ri[e] + csub(mont_reduce(ar, M.q, M.qinv), M.q), M.q);
                    if constexpr (WY) S[e] = csub(S[e] + csub(mont_reduce(ab, M.q, M.qinv), M.q), M.q);
                }
            }
        };
        if (use_y) mac(std::true_type{});
        else mac(std::false_type{});
    }

    // a_hat = INTT(R[0]); b = b_poly + INTT(S)
    uint32_t* out = A.out + (A.roles.qout ? A.roles.qout[qi] : (uint64_t)qi * 2 * KN) + (size_t)k * N;
    intt_v<LOGN, V>(R, t, rb, tw, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
#pragma unroll
    for (int e = 0; e < V; e++) out[t + e * T] = csub(R[e], M.q);
    lds_barrier();  // every thread is past its last read of rb
    intt_v<LOGN, V>(S, t, rb, tw, M, inv.c[k], inv.cp[k], inv.cw[k], inv.cwp[k]);
#pragma unroll
    for (int e = 0; e < V; e++) {
        const uint32_t pos = t + e * T;
        uint32_t v = csub(S[e], M.q);
        if (pos < gamma) {
            // b_poly coefficient: n^-1 * (sum of the column's bhat) mod q_k.
            // The sum is below 2^40: (s 2^-32) (n^-1 2^32) = s n^-1.
            uint64_t s = 0;
            for (uint32_t pb = 0; pb < A.nparts; pb++)
                s += A.bpart[(((size_t)qi * A.nparts + pb) * gamma + pos) * 2 + k];
            const uint32_t a = csub(mont_reduce(s, M.q, M.qinv), M.q);
            v = csub(v + csub(shoup_mul(a, A.ninvm[k], A.ninvmp[k], M.q), M.q), M.q);
        }
        out[KN + pos] = v;
    }
}

// Raises the kernel's dynamic LDS limit once per instantiation (a function-
// local static is initialised exactly once, also under concurrent callers).
template <int LOGN, int ELL, bool LAZY, bool REC, bool SGN, int V>
static hipError_t chain_crt_lds_limit() {
    constexpr size_t shmem = ChainCrtShape<LOGN, ELL, REC, V>::kBytes;
    static const hipError_t e =
        shmem > 64 * 1024 ? hipFuncSetAttribute((const void*)k_pack_chain_crt<LOGN, ELL, LAZY, REC, SGN, V>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)
                          : hipSuccess;
    return e;
}

template <int LOGN, int ELL, bool LAZY, bool REC, bool SGN, int V>
static hipError_t chain_crt_launch_v(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                     hipStream_t st, LaunchEv ev) {
    const hipError_t e = chain_crt_lds_limit<LOGN, ELL, LAZY, REC, SGN, V>();
    if (e != hipSuccess) return e;
    auto kern = k_pack_chain_crt<LOGN, ELL, LAZY, REC, SGN, V>;
    constexpr size_t shmem = ChainCrtShape<LOGN, ELL, REC, V>::kBytes;
    hipExtLaunchKernelGGL(kern, dim3(Q), dim3(2 * ((1 << LOGN) / V)), shmem, st, ev.start, ev.stop, 0, A, C, inv);
    return hipGetLastError();
}

// Values per thread: four by default (2 N/4 threads per query, four waves
// per SIMD at d = 2048), where eight (the radix-8 layout) leave two waves per
// SIMD to hide each step's LDS exchanges and barriers; CtxDev.chain_v selects
// (the tuner measures both).
constexpr int kChainV = 4;

template <int LOGN, int ELL, bool LAZY, bool REC, bool SGN = false>
static hipError_t chain_crt_launch(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                   hipStream_t st, LaunchEv ev) {
    if (C.chain_v == 8) return chain_crt_launch_v<LOGN, ELL, LAZY, REC, SGN, 8>(A, Q, C, inv, st, ev);
    return chain_crt_launch_v<LOGN, ELL, LAZY, REC, SGN, kChainV>(A, Q, C, inv, st, ev);
}

// The SGN variant's bounds (signed_digits_ok): R < 2q enters the recovered
// digit, r < q and S < 2q are folded into the multiply-accumulates, and
// q0 < 2 q1 (a residue mod q0 is reduced mod q1 by one subtraction).
static bool chain_signed_ok(const CtxDev& C, int logn, int ell, bool rec) {
    if (C.m[0].q >= 2ull * C.m[1].q) return false;
    if (!C.small_digits) return false;  // the signed kernel takes the gadget digits unreduced
    // the signed inverse transforms hold sums below 8 q in int32
    if (8ull * C.m[0].q >= (1ull << 31) || 8ull * C.m[1].q >= (1ull << 31)) return false;
    for (int k = 0; k < 2; k++) {
        const uint64_t q = C.m[k].q;
        if (!signed_digits_ok(q, C.small_digits ? C.gadget_mask : q - 1, logn, rec ? ell - 1 : ell, rec ? 2 * q : 0,
                              2 * q))
            return false;
    }
    return true;
}

template <int LOGN, int ELL>
static hipError_t chain_crt_t(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv, hipStream_t st,
                              LaunchEv ev) {
    const uint64_t qmax = C.m[0].q > C.m[1].q ? C.m[0].q : C.m[1].q;
    // the recovered digit needs the composite value to fit ell digits: q <= z^ell
    const bool rec = (uint64_t)C.gadget_log * ELL >= 64 || C.q <= (1ull << (C.gadget_log * ELL));
    if ((uint64_t)ELL * 4 * qmax < (1ull << 32)) {
        if (chain_signed_ok(C, LOGN, ELL, rec)) {
            if (rec) return chain_crt_launch<LOGN, ELL, true, true, true>(A, Q, C, inv, st, ev);
            return chain_crt_launch<LOGN, ELL, true, false, true>(A, Q, C, inv, st, ev);
        }
        if (rec) return chain_crt_launch<LOGN, ELL, true, true>(A, Q, C, inv, st, ev);
        return chain_crt_launch<LOGN, ELL, true, false>(A, Q, C, inv, st, ev);
    }
    if (rec) return chain_crt_launch<LOGN, ELL, false, true>(A, Q, C, inv, st, ev);
    return chain_crt_launch<LOGN, ELL, false, false>(A, Q, C, inv, st, ev);
}

template <int LOGN>
static hipError_t chain_crt_ell(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                hipStream_t st, LaunchEv ev) {
    switch (C.ell) {
        case 1: return chain_crt_t<LOGN, 1>(A, Q, C, inv, st, ev);
        case 2: return chain_crt_t<LOGN, 2>(A, Q, C, inv, st, ev);
        case 3: return chain_crt_t<LOGN, 3>(A, Q, C, inv, st, ev);
        default: return hipErrorInvalidValue;
    }
}

bool pack_chain_crt_supported(const CtxDev& C) {
    if (C.K != 2 || C.ell < 1 || C.ell > 3 || C.logn < 8 || C.logn > 11) return false;
    // every modulus must leave room for the lazy ranges: 4q < 2^32
    if (C.m[0].q >= (1u << 30) || C.m[1].q >= (1u << 30)) return false;
    return ChainCrtShape<11, 3, false, kChainV>::kBytes <= 160 * 1024 && ChainCrtShape<11, 3, false, 8>::kBytes <= 160 * 1024;
}

hipError_t launch_pack_chain_crt(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                 hipStream_t st, LaunchEv ev) {
    if (!Q) return hipSuccess;
    switch (C.logn) {
        case 8: return chain_crt_ell<8>(A, Q, C, inv, st, ev);
        case 9: return chain_crt_ell<9>(A, Q, C, inv, st, ev);
        case 10: return chain_crt_ell<10>(A, Q, C, inv, st, ev);
        case 11: return chain_crt_ell<11>(A, Q, C, inv, st, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
