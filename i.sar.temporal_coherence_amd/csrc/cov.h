/* SHP-masked covariance and coherence matrix of one pixel (tcoh.cl:567-
 * 639): C_ij = sum over accepted candidates of s_i conj(s_j), then
 * T_ij = C_ij / sqrt(C_ii C_jj) with the optional bias correction of the
 * magnitudes, and the pair coherences |T_ij|. T (i < j) goes to LDS,
 * packed row by row ((0,1),(0,2)...,(1,2)...), for the linking stage; the
 * pair planes go to global memory. Two arithmetic forms:
 *
 * plain (cov_plain_group): FP32 on the vector ALUs, one pixel per group
 *   of G lanes. The upper triangle (with diagonal) is cut into NBLK x NBLK
 *   blocks of B x B; lane l owns upper block l and keeps its B*B complex
 *   sums in registers; per accepted candidate (dy-major, the reference's
 *   summation order) it loads the B row and B column samples it needs.
 *   EMI takes this form with FP64 sums, whatever the precision setting:
 *   |T|^-1 o T amplifies the rounding of FP32 sums (up to 9e-5 in gamma on
 *   near-singular |T|), and of the split forms' representation likewise.
 *
 * mixed (cov_mfma_band): the whole wavefront on one pixel at a time, as
 *   real Gram products. With Re, Im the N x L matrices of the L accepted
 *   candidates, C_re = Re Re^T + Im Im^T and C_im = Im Re^T - Re Im^T.
 *   Each sample is split as hi + lo (FP16: x = hi + lo + O(2^-22 x), each
 *   pixel under its own power-of-two scale, which the gather turns into
 *   the window's common one, exactly, by the candidate's exponent below
 *   the window's largest; BF16: O(2^-18 x), unscaled) and every product
 *   is accumulated in FP32 from hi.hi +
 *   hi.lo + lo.hi with v_mfma_f32_16x16x32_{f16,bf16}: C_re for the upper
 *   16 x 16 date blocks only; C_im up to 32 date slots as X - X^T from
 *   X = Im Re^T over all blocks (three products a block pair instead of
 *   six), above that directly. The accepted candidates are compacted into
 *   LDS lists padded to a multiple of 32 with a zero pixel; A and B
 *   operands share one lane layout, so the candidate order inside a
 *   32-candidate chunk does not matter. Dates are padded to NB with zeros.
 */
#ifndef TCOH_COV_H
#define TCOH_COV_H

#include "common.h"

/* Normalised and (optionally) debiased T_ij from the sums (tcoh.cl:610-
   636); den = sqrt(C_ii) sqrt(C_jj). */
__device__ __forceinline__ float2 coherence_entry(float cre, float cim,
                                                  float den, int bias,
                                                  int looks)
{
    float re = 0.0f, im = 0.0f;

    if (den > 0.0f) {
        re = cre / den;
        im = cim / den;
        if (bias && looks > 1) {
            const float m2 = re * re + im * im;
            const float m2c = (looks * m2 - 1.0f) / (looks - 1.0f);
            const float f = (m2 > 0.0f && m2c > 0.0f) ? sqrt_fast(m2c / m2) : 0.0f;

            re *= f;
            im *= f;
        }
    }
    return make_float2(re, im);
}

/* The same in FP64 (EMI), rounded to float at the end, from
   s = 1 / sqrt(C_ii C_jj) (0 where a diagonal sum is 0), with the bias
   factor as one quotient: m2c / m2 = (L m2 - 1) /
   ((L - 1) m2), and m2c > 0 exactly where that quotient is. */
__device__ __forceinline__ float2 coherence_entry_scaled(double cre, double cim, double s,
                                                         int bias, int looks)
{
    double re = cre * s, im = cim * s;

    if (bias && looks > 1) {
        const double m2 = re * re + im * im;
        const double q = m2 > 0.0 ? (looks * m2 - 1.0) / ((looks - 1) * m2) : 0.0;
        const double f = q > 0.0 ? sqrt(q) : 0.0;

        re *= f;
        im *= f;
    }
    return make_float2((float)re, (float)im);
}

/* Pair plane of T_ij, or -1 if (i, j) is not an output pair. */
__device__ __forceinline__ int pair_plane(const KParams &kp, int i, int j)
{
    if (!kp.npairs)
        return -1;
    return kp.pairs_all ? upper_index(kp.n, i, j) : (j == i + 1 ? i : -1);
}

template <int G> struct PlainShape {
    static constexpr int NBLK = G == 8 ? 3 : (G == 16 ? 5 : 10);
};

/* Plain form, one pixel per group, sums in ACC: float (the reference's
   arithmetic) or double (EMI, whose |T|^-1 o T amplifies the rounding of
   FP32 sums). diag: LDS scratch of N ACC values (the diagonal sums; in
   FP64 their reciprocal square roots); mw: of the pixel's kp.words mask
   words. Called by every lane of the wavefront. */
template <int N, int G, typename ACC>
__device__ __forceinline__ void cov_plain_group(const KParams &kp,
                                                const float2 *__restrict__ slc,
                                                const uint32_t *__restrict__ mask,
                                                int p, int cnt, bool est,
                                                bool live, int lg, ACC *diag,
                                                uint32_t *mw, float2 *tl,
                                                float *__restrict__ pairs)
{
    constexpr int NBLK = PlainShape<G>::NBLK;
    constexpr int B = (N + NBLK - 1) / NBLK;
    const size_t plane = (size_t)kp.rows * kp.cols;
    const int y = p / kp.cols, x = p - (p / kp.cols) * kp.cols;
    int bi = 0, bj = 0;
    ACC cr[B][B], ci[B][B];

    /* Upper block number lg -> (bi, bj), bi <= bj. */
    {
        int t = lg;

        while (bi < NBLK && t >= NBLK - bi) {
            t -= NBLK - bi;
            bi++;
        }
        bj = bi + t;
    }
    const bool own = lg < NBLK * (NBLK + 1) / 2;
    /* block bases clamped so that every date a lane reads is below N (the
       last blocks of a row or column then overlap the ones before; both
       lanes sum the shared entries in the same order) */
    const int r0 = min(bi * B, N - B), c0 = min(bj * B, N - B);

#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++)
            cr[r][c] = ci[r][c] = (ACC)0;

    /* The pixel's mask words go to LDS first: read from global memory
       inside the walk, each word's load made the trip wait for every load
       in flight, the samples' included. */
    if (live)
        for (int w = lg; w < kp.words; w += G)
            mw[w] = mask[(size_t)p * kp.words + w];
    wave_sync();
    if (est && own) {
        /* Accepted candidates in window raster order (the reference's
           summation order). Three sample buffers rotate: the loads of each
           candidate are issued two candidates before it is summed, so that
           their latency overlaps the multiply-adds (the FP64 form leaves
           room for two wavefronts per SIMD at most, which hide little). */
        const float inv_wr = 1.0f / (float)kp.wr;
        uint32_t bits = mw[0];
        int wi = 0;
        /* bit index of the next accepted candidate, -1 past the last */
        auto next = [&]() -> int {
            while (bits == 0u) {
                if (++wi >= kp.words)
                    return -1;
                bits = mw[wi];
            }
            const int t = wi * 32 + __builtin_ctz(bits);

            bits &= bits - 1u;
            return t;
        };
        /* samples of candidate t (loads issued in every lane, from the
           window's first candidate past the last) */
        auto fetch = [&](int t, float2 *fr, float2 *fc) {
            const int tt = t < 0 ? 0 : t;
            /* exact: (tt + 0.5) / wr is at least 0.5 / wr from an integer
               and tt, wr <= TCOH_MAX_WINDOW */
            const int dy = (int)(((float)tt + 0.5f) * inv_wr);
            const float2 *s = slc + ((size_t)(y + dy) * kp.w + x + (tt - dy * kp.wr)) * N;

#pragma unroll
            for (int r = 0; r < B; r++) {
                fr[r] = s[r0 + r];
                fc[r] = s[c0 + r];
            }
            /* keep the loads ahead of the multiply-adds that follow */
            __builtin_amdgcn_sched_barrier(0);
        };
        auto sum = [&](const float2 *sr, const float2 *sc) {
#pragma unroll
            for (int r = 0; r < B; r++)
#pragma unroll
                for (int c = 0; c < B; c++) {
                    const ACC ar = sr[r].x, ai = sr[r].y, br = sc[c].x, bi_ = sc[c].y;

                    if constexpr (sizeof(ACC) == 8) {
                        /* each FP32 product is exact in FP64: one
                           rounding a term */
                        cr[r][c] = fma(ai, bi_, fma(ar, br, cr[r][c]));
                        ci[r][c] = fma(-ar, bi_, fma(ai, br, ci[r][c]));
                    }
                    else {
                        cr[r][c] += ar * br + ai * bi_;
                        ci[r][c] += ai * br - ar * bi_;
                    }
                }
        };
        float2 ar_[B], ac_[B], br_[B], bc_[B], cr_[B], cc_[B];
        int ta = next(), tb, tc;

        fetch(ta, ar_, ac_);
        tb = next();
        fetch(tb, br_, bc_);
        while (ta >= 0) {
            tc = next();
            fetch(tc, cr_, cc_);
            sum(ar_, ac_);
            if (tb < 0)
                break;
            ta = next();
            fetch(ta, ar_, ac_);
            sum(br_, bc_);
            if (tc < 0)
                break;
            tb = next();
            fetch(tb, br_, bc_);
            sum(cr_, cc_);
        }
    }
    if (own && bi == bj) {
#pragma unroll
        for (int r = 0; r < B; r++)
            if (r0 + r < N)
                /* FP64 (EMI): the reciprocal square root, so that each
                   entry takes two products instead of two divisions */
                diag[r0 + r] = sizeof(ACC) == 8
                                   ? (cr[r][r] > (ACC)0 ? (ACC)1 / sqrt(cr[r][r]) : (ACC)0)
                                   : cr[r][r];
    }
    wave_sync();
    if (!live || !own)
        return;
#pragma unroll
    for (int r = 0; r < B; r++) {
        const int i = r0 + r;

#pragma unroll
        for (int c = 0; c < B; c++) {
            const int j = c0 + c;

            if (i >= N || j >= N || j <= i)
                continue;
            const int pidx = pair_plane(kp, i, j);

            if (!est) {
                if (pidx >= 0)
                    pairs[(size_t)pidx * plane + p] = __builtin_nanf("");
                continue;
            }
            float2 t;

            if constexpr (sizeof(ACC) == 8)
                t = coherence_entry_scaled(cr[r][c], ci[r][c], diag[i] * diag[j], kp.bias, cnt);
            else
                t = coherence_entry(cr[r][c], ci[r][c], sqrt_fast(diag[i]) * sqrt_fast(diag[j]),
                                    kp.bias, cnt);

            tl[upper_index(N, i, j)] = t;
            if (pidx >= 0)
                pairs[(size_t)pidx * plane + p] = sqrt_fast(t.x * t.x + t.y * t.y);
        }
    }
}

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

template <int BF16>
__device__ __forceinline__ float4_t mfma(uint4_t a, uint4_t b, float4_t c)
{
    if (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
            0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(
            __builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b), c, 0,
            0, 0);
}

/* Operands of one 16-date block for the lane's 8 candidates: low halves
   (hi parts) and high halves (lo parts) of the re and im words. */
struct Opnd {
    uint4_t rh, rl, ih, il;
};

__device__ __forceinline__ Opnd make_opnd(const uint2 (&v)[8])
{
    Opnd o;

#pragma unroll
    for (int q = 0; q < 4; q++) {
        o.rh[q] = __builtin_amdgcn_perm(v[2 * q + 1].x, v[2 * q].x, 0x05040100u);
        o.rl[q] = __builtin_amdgcn_perm(v[2 * q + 1].x, v[2 * q].x, 0x07060302u);
        o.ih[q] = __builtin_amdgcn_perm(v[2 * q + 1].y, v[2 * q].y, 0x05040100u);
        o.il[q] = __builtin_amdgcn_perm(v[2 * q + 1].y, v[2 * q].y, 0x07060302u);
    }
    return o;
}

/* Candidate lists of the mixed form: the byte offset of every accepted
   candidate's sample row from the band's first window row (the base of a
   range-checked buffer resource over the band's rows), padded to a
   multiple of 32 with LIST_PAD, beyond every band: its loads return zero
   without touching memory; in the FP16 form a parallel list holds each
   candidate's
   rescaling factor 2^-(its exponent below the window's largest), as a pair
   of FP16 values. */
#define LIST_SHIFT_MAX 31
#define LIST_PAD 0xFFFF0000u
/* Third descriptor word of a raw buffer resource on gfx9 parts (32-bit
   data format, no swizzle); loads past the range return zero. */
#define RAW_BUFFER_WORD3 0x00020000
#define F16_ONE_PAIR 0x3C003C00u

/* FP16 bits of 2^-s, 0 <= s <= 31 (subnormal from s = 15, zero above
   24), in both halves of a word. */
__device__ __forceinline__ uint32_t f16_pow2_neg_pair(int s)
{
    const uint32_t b = s <= 14 ? (uint32_t)(15 - s) << 10
                               : (s <= 24 ? 1u << (24 - s) : 0u);

    return b | (b << 16);
}

/* Both FP16 halves of w times those of f (v_pk_mul_f16: exact for a power
   of two unless the product is subnormal, then rounded to nearest). */
__device__ __forceinline__ uint32_t pk_mul_f16(uint32_t w, uint32_t f)
{
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(half2_t, w) *
                                            __builtin_bit_cast(half2_t, f));
}
__host__ __device__ __forceinline__ int mfma_list_stride(int kw)
{
    return (kw + 31) & ~31;
}

/* Bytes of the second list area: the FP16 form's factors, or (BF16, up to
   32 date slots) the transposition tiles of the epilogue, one 16 x 16
   float tile per upper block pair, whichever is larger. */
__host__ __device__ __forceinline__ int mfma_area2_bytes(int ppw, int kw, int nb)
{
    const int fac = ppw * mfma_list_stride(kw) * 4;
    const int nbk = nb / 16;
    const int xt = nb <= 32 ? nbk * (nbk + 1) / 2 * 1024 : 0;

    return fac > xt ? fac : xt;
}

/* LDS bytes of the mixed form for ppw pixels: lists, the second area,
   the diagonal of C, the pixels' mask words, and a per-pixel table
   (chunks, looks). */
__host__ __device__ __forceinline__ int mfma_scratch_bytes(int ppw, int kw,
                                                           int words, int nb)
{
    return ppw * mfma_list_stride(kw) * 4 + mfma_area2_bytes(ppw, kw, nb) + nb * 4 +
           ppw * words * 4 + ppw * 8;
}

/* Mixed form over the group's pixels x0 .. x0+PPW-1 of output row y, the
   whole wavefront on one pixel at a time. The mask words of all pixels are
   staged once and every candidate list is built (mbcnt compaction) before
   the first gather, so one gather stream, DEPTH chunks ahead of the
   MFMAs, runs through all pixels without draining at pixel boundaries.
   C_re = Re Re^T + Im Im^T and C_im = Im Re^T - Re Im^T accumulate in
   registers for the upper 16 x 16 blocks only (the sign by flipping the
   16-bit sign bits of Re); the epilogue normalises them from registers. */
template <int N, int NB, int PPW, int BF16>
__device__ __forceinline__ void cov_mfma_band(const KParams &kp,
                                              const uint2 *__restrict__ hl,
                                              const short *__restrict__ ex,
                                              const uint32_t *__restrict__ mask,
                                              const int *__restrict__ count,
                                              int y, int x0, int row0, int nrows,
                                              unsigned char *scratch,
                                              float2 *tl,
                                              float *__restrict__ pairs)
{
    constexpr int NBK = NB / 16;             /* 16-date blocks */
    constexpr int NUP = NBK * (NBK + 1) / 2; /* upper block pairs */
    constexpr int NT = N * (N - 1) / 2;
    /* chunks in flight: at 64 date slots the buffers compete with the
       accumulators for registers, and from 48 dates on a single buffer
       (no prefetch: fewer spills at two wavefronts per SIMD) is faster
       (N=52, 60, 64: 4, 7, 2 % on 512^2-768^2 frames; N=45: 3 % slower) */
    constexpr int DEPTH = NBK <= 2 ? 2 : (N >= 48 ? 0 : 1);
    const int lane = threadIdx.x;
    const int kw = kp.wa * kp.wr, nw = kp.words;
    const int lstride = mfma_list_stride(kw);
    /* Up to 32 date slots: C_im = X - X^T from X = Im Re^T over all
       block pairs (a, b), which takes 3 split products per pair instead
       of 6 per upper pair (NBK = 2: 12 MFMAs a chunk instead of 18) and
       no sign flips; the epilogue transposes X(b, a) through LDS: BF16 in
       the second list area (it has no factors), FP16 in the pixel's own T
       area before T is written, where that holds the tiles (N >= 29). */
    constexpr int XT_BYTES = NBK * (NBK + 1) / 2 * 1024;
    constexpr bool XT = NBK <= 2 && (BF16 || NT * 8 - 8 >= XT_BYTES);
    int *list = reinterpret_cast<int *>(scratch);
    uint32_t *fac = reinterpret_cast<uint32_t *>(scratch + PPW * lstride * 4);
    float *xtl = reinterpret_cast<float *>(fac);
    float *diag = reinterpret_cast<float *>(scratch + PPW * lstride * 4 +
                                            mfma_area2_bytes(PPW, kw, NB));
    uint32_t *mw = reinterpret_cast<uint32_t *>(diag + NB);
    /* chunks and looks of pixel q at ptab[q], ptab[PPW + q]: read back with
       a uniform address when q is only known at run time (a register
       array indexed so would become a scratch-memory table, whose loads
       would wait behind the gathers in flight) */
    int *ptab = reinterpret_cast<int *>(mw + PPW * nw);
    const size_t plane = (size_t)kp.rows * kp.cols;
    const int p0 = y * kp.cols + x0;
    const int npix = min(PPW, kp.cols - x0);
    int cnt[PPW], nch[PPW];

#pragma unroll
    for (int q = 0; q < PPW; q++)
        cnt[q] = q < npix ? __builtin_amdgcn_readfirstlane(count[p0 + q]) : -1;
    for (int t = lane; t < npix * nw; t += WAVE)
        mw[t] = mask[(size_t)p0 * nw + t];
    wave_sync();

    const float inv_wr = 1.0f / (float)kp.wr;
    int total = 0;

#pragma unroll
    for (int q = 0; q < PPW; q++) {
        nch[q] = 0;
        if (cnt[q] < kp.min_looks) {
            if (q < npix && kp.npairs)
                for (int t = lane; t < kp.npairs; t += WAVE)
                    pairs[(size_t)t * plane + p0 + q] = __builtin_nanf("");
            continue;
        }
        int *lq = list + q * lstride;
        const int origin = y * kp.w + x0 + q;
        /* the band's first window row: list offsets count from it */
        const int band_px = row0 * kp.w;
        int base = 0, emax = SPLIT_EX_ZERO;

        for (int it = 0; it * 64 < kw; it++) {
            /* The words are uniform: read them into scalar registers, so
               that the counts and every branch on them stay uniform. */
            unsigned long long m64 =
                (unsigned)__builtin_amdgcn_readfirstlane((int)mw[q * nw + 2 * it]);

            if (2 * it + 1 < nw)
                m64 |= (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(
                           (int)mw[q * nw + 2 * it + 1])
                       << 32;
            if (kw - it * 64 < 64)
                m64 &= (1ull << (kw - it * 64)) - 1ull;
            const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(m64 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m64, 0u));

            if ((m64 >> lane) & 1ull) {
                /* b < 1024: (b + 0.5) / wr is at least 0.5 / wr away from
                   an integer, far more than the float product's error. */
                const int b = it * 64 + lane;
                const int dy = (int)(((float)b + 0.5f) * inv_wr);

                const int cand = origin + dy * kp.w + (b - dy * kp.wr);

                lq[base + rank] = (int)((uint32_t)(cand - band_px) * (uint32_t)(NB * 8));
                if (!BF16) {
                    /* the exponent beside it, for the factor below */
                    const int e = ex[cand];

                    fac[q * lstride + base + rank] = (uint32_t)e;
                    emax = max(emax, e);
                }
            }
            base += __popcll(m64);
        }
        const int lpad = (base + 31) & ~31;

        if (!BF16) {
            /* FP16: each candidate's factor 2^-(its exponent below the
               window's largest accepted one), by which step() rescales its
               hi and lo parts, in place of the exponent */
            uint32_t *fq = fac + q * lstride;

            emax = gmaxi<WAVE>(emax);
            wave_sync();
            for (int t = lane; t < lpad; t += WAVE)
                fq[t] = t < base ? f16_pow2_neg_pair(min(emax - (int)fq[t], LIST_SHIFT_MAX))
                                 : F16_ONE_PAIR;
        }
        for (int t = base + lane; t < lpad; t += WAVE)
            lq[t] = (int)LIST_PAD;
        nch[q] = lpad >> 5;
        total += nch[q];
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < PPW; q++) {
            ptab[q] = nch[q];
            ptab[PPW + q] = cnt[q];
        }
    }
    wave_sync();
    if (total == 0)
        return;

    const int i16 = lane & 15, g = lane >> 4;
    /* xim[a * NBK + b] = X(a, b) (XT only; cim is then formed in the
       epilogue) */
    constexpr int NX = XT ? NBK * NBK : 1;
    float4_t cre[NUP], cim[NUP], xim[NX];

#pragma unroll
    for (int t = 0; t < NUP; t++)
        cre[t] = cim[t] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < NX; t++)
        xim[t] = float4_t{0.f, 0.f, 0.f, 0.f};

    /* Pixel and chunk of flat chunk c (pixels without estimate have no
       chunks). */
    auto locate = [&](int c, int &q, int &off) {
        q = 0;
        off = c;
#pragma unroll
        for (int k = 0; k < PPW - 1; k++)
            if (q == k && off >= nch[k]) {
                off -= nch[k];
                q = k + 1;
            }
    };
    /* Samples of flat chunk c for this lane: dates bk*16 + i16 of the 8
       candidates of its group g. hl rows hold NB date slots (zero past N),
       so every load is unconditional. */
    /* The band's sample rows (window rows row0 .. row0 + nrows + wa - 2)
       as a range-checked buffer: a list offset is one 32-bit add from the
       lane's date, and the pads (LIST_PAD) load zeros. The host keeps the
       span within LIST_SPAN_MAX. */
    const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint2 *>(hl + (size_t)row0 * kp.w * NB), 0,
        (int)((uint32_t)(nrows + kp.wa - 1) * (uint32_t)kp.w * (uint32_t)(NB * 8)),
        RAW_BUFFER_WORD3);
    auto fetch = [&](int c, uint2(&v)[NBK][8]) {
        int q, off;

        locate(c, q, off);
        /* two 16-byte LDS reads (list rows are 128-byte aligned) */
        const int4 *lp = reinterpret_cast<const int4 *>(list + q * lstride + off * 32 + g * 8);
        const int4 i0 = lp[0], i1 = lp[1];
        const int idx[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};

#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int o = (int)((uint32_t)idx[e] + (uint32_t)(i16 * 8));

#pragma unroll
            for (int bk = 0; bk < NBK; bk++) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b64(rows_rsrc, o + bk * 128, 0, 0);

                v[bk][e] = make_uint2(w[0], w[1]);
            }
        }
    };

    int q, left;

    locate(0, q, left);
    left = __builtin_amdgcn_readfirstlane(ptab[q]);

    /* One chunk: MFMAs on use (chunk c), and at the pixel's last chunk the
       epilogue. */
    auto step = [&](int c, const uint2(&use)[NBK][8]) {
        Opnd op[NBK];
        uint4_t nrh[NBK], nrl[NBK];
        uint2 sv[NBK][8];

        if (!BF16) {
            /* the window's common scale: each candidate's hi and lo times
               2^-(its exponent below the largest), from the list entries
               of this chunk (read again here: the gathers were issued
               chunks ago) */
            int cq, off;

            locate(c, cq, off);
            const uint4 *fp = reinterpret_cast<const uint4 *>(fac + cq * lstride + off * 32 + g * 8);
            const uint4 f0 = fp[0], f1 = fp[1];
            const uint32_t fe[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};

#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint32_t f = fe[e];

#pragma unroll
                for (int bk = 0; bk < NBK; bk++)
                    sv[bk][e] = make_uint2(pk_mul_f16(use[bk][e].x, f),
                                           pk_mul_f16(use[bk][e].y, f));
            }
        }
        else {
#pragma unroll
            for (int bk = 0; bk < NBK; bk++)
#pragma unroll
                for (int e = 0; e < 8; e++)
                    sv[bk][e] = use[bk][e];
        }

#pragma unroll
        for (int bk = 0; bk < NBK; bk++) {
            op[bk] = make_opnd(sv[bk]);
            if (!XT) {
                nrh[bk] = op[bk].rh ^ 0x80008000u;
                nrl[bk] = op[bk].rl ^ 0x80008000u;
            }
        }
        /* hi.hi + hi.lo + lo.hi of each product, the accumulators of a
           step independent of each other. */
#define TCOH_UPPER(STMT)                                                       \
    {                                                                          \
        int t = 0;                                                             \
        _Pragma("unroll") for (int a = 0; a < NBK; a++)                        \
            _Pragma("unroll") for (int b = a; b < NBK; b++, t++) { STMT; }     \
    }
#define TCOH_ALL(STMT)                                                         \
    _Pragma("unroll") for (int a = 0; a < NBK; a++)                            \
        _Pragma("unroll") for (int b = 0; b < NBK; b++) { STMT; }
        if constexpr (XT) {
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rh, op[b].rh, cre[t]))
            TCOH_ALL(xim[a * NBK + b] = mfma<BF16>(op[a].ih, op[b].rh, xim[a * NBK + b]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].ih, op[b].ih, cre[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rh, op[b].rl, cre[t]))
            TCOH_ALL(xim[a * NBK + b] = mfma<BF16>(op[a].ih, op[b].rl, xim[a * NBK + b]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].ih, op[b].il, cre[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rl, op[b].rh, cre[t]))
            TCOH_ALL(xim[a * NBK + b] = mfma<BF16>(op[a].il, op[b].rh, xim[a * NBK + b]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].il, op[b].ih, cre[t]))
        }
        else {
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rh, op[b].rh, cre[t]);
                       cim[t] = mfma<BF16>(op[a].ih, op[b].rh, cim[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rh, op[b].rl, cre[t]);
                       cim[t] = mfma<BF16>(op[a].ih, op[b].rl, cim[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].rl, op[b].rh, cre[t]);
                       cim[t] = mfma<BF16>(op[a].il, op[b].rh, cim[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].ih, op[b].ih, cre[t]);
                       cim[t] = mfma<BF16>(nrh[a], op[b].ih, cim[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].ih, op[b].il, cre[t]);
                       cim[t] = mfma<BF16>(nrh[a], op[b].il, cim[t]))
            TCOH_UPPER(cre[t] = mfma<BF16>(op[a].il, op[b].ih, cre[t]);
                       cim[t] = mfma<BF16>(nrl[a], op[b].ih, cim[t]))
        }
#undef TCOH_ALL
#undef TCOH_UPPER

        if (--left > 0)
            return;

        /* Last chunk of pixel q: T from the accumulators, lane (i16, g)
           holding D[4g + r][i16] of each tile. */
        const int looks = __builtin_amdgcn_readfirstlane(ptab[PPW + q]);
        float2 *tq = tl + q * NT;

        {
            int t = 0;

#pragma unroll
            for (int a = 0; a < NBK; a++)
#pragma unroll
                for (int b = a; b < NBK; b++, t++)
                    if (a == b && (i16 >> 2) == g) {
                        const int r = i16 & 3;
                        const float v = r == 0 ? cre[t][0]
                                        : r == 1 ? cre[t][1]
                                        : r == 2 ? cre[t][2] : cre[t][3];

                        diag[a * 16 + i16] = v;
                    }
        }
        /* the transposition area (16-byte aligned) */
        float *xa = xtl;

        if (!BF16) {
            const uintptr_t u = reinterpret_cast<uintptr_t>(tq);

            xa = reinterpret_cast<float *>((u + 15) & ~static_cast<uintptr_t>(15));
        }
        if constexpr (XT) {
            /* X(b, a) of each upper pair (a, b) into tile t of the area,
               column-major (element (i, j) at j * 16 + i) */
            int t = 0;

#pragma unroll
            for (int a = 0; a < NBK; a++)
#pragma unroll
                for (int b = a; b < NBK; b++, t++)
                    *reinterpret_cast<float4_t *>(xa + t * 256 + i16 * 16 + 4 * g) =
                        xim[b * NBK + a];
        }
        wave_sync();
        if constexpr (XT) {
            /* C_im(a, b)[i][j] = X(a, b)[i][j] - X(b, a)[j][i], with
               i = 4g + r and j = i16 */
            int t = 0;

#pragma unroll
            for (int a = 0; a < NBK; a++)
#pragma unroll
                for (int b = a; b < NBK; b++, t++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        cim[t][r] = xim[a * NBK + b][r] - xa[t * 256 + (4 * g + r) * 16 + i16];
            /* FP16: the area is the T being written below */
            if (!BF16)
                wave_sync();
        }
        float rsr[NBK][4], rsc[NBK];
        /* FP16: the diagonal sums squares of FP16 values (at least 2^-48
           when not zero), never an FP32 denormal, so the hardware rsq
           needs no range scaling; BF16 samples are unscaled FP32 values */
        auto rsq = [](float v) { return BF16 ? rsqrtf(v) : __builtin_amdgcn_rsqf(v); };

#pragma unroll
        for (int a = 0; a < NBK; a++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float dv = diag[a * 16 + 4 * g + r];

                rsr[a][r] = dv > 0.0f ? rsq(dv) : 0.0f;
            }
            const float dv = diag[a * 16 + i16];

            rsc[a] = dv > 0.0f ? rsq(dv) : 0.0f;
        }
        const bool bias = kp.bias && looks > 1;
        const float fl = (float)looks;
        /* f^2 = m2c / m2 = (L m2 - 1) / ((L - 1) m2) = L/(L-1) - 1/((L-1) m2) */
        const float inv_l1 = bias ? 1.0f / (fl - 1.0f) : 0.0f;
        const float fa = fl * inv_l1;
        int t = 0;

#pragma unroll
        for (int a = 0; a < NBK; a++)
#pragma unroll
            for (int b = a; b < NBK; b++, t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = a * 16 + 4 * g + r, j = b * 16 + i16;
                    const float sc = rsr[a][r] * rsc[b];
                    float re = cre[t][r] * sc, im = cim[t][r] * sc;

                    if (bias) {
                        /* tcoh.cl:620-633: the magnitude scaled by
                           sqrt(m2c / m2), zero where m2c <= 0; the same
                           threshold L m2 = 1 decided on f^2 (m2 = 0: the
                           hardware reciprocal is +inf, f^2 -inf, f 0) */
                        const float m2 = re * re + im * im;
                        const float f2 = fmaf(-inv_l1, __builtin_amdgcn_rcpf(m2), fa);
                        const float f = f2 > 0.0f ? sqrt_fast(f2) : 0.0f;

                        re *= f;
                        im *= f;
                    }
                    if (i < j && j < N)
                        tq[upper_index(N, i, j)] = make_float2(re, im);
                }
#pragma unroll
        for (int u = 0; u < NUP; u++) {
            cre[u] = float4_t{0.f, 0.f, 0.f, 0.f};
            if (!XT) /* (XT: cim is formed anew in each epilogue) */
                cim[u] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int u = 0; u < NX; u++)
            xim[u] = float4_t{0.f, 0.f, 0.f, 0.f};
        if (c + 1 < total) {
            int off;

            locate(c + 1, q, off);
            left = __builtin_amdgcn_readfirstlane(ptab[q]) - off;
        }
        wave_sync();
    };

    /* Buffers named statically (no register copies) and fetched on every
       path (past the end the last chunk again), so that each step waits
       for the loads of its own chunk only. */
    uint2 b0[NBK][8], b1[NBK][8], b2[NBK][8];
    const int last = total - 1;

    if (DEPTH == 2) {
        fetch(0, b0);
        fetch(min(1, last), b1);
        for (int c = 0; c < total; c += 3) {
            fetch(min(c + 2, last), b2);
            step(c, b0);
            if (c + 1 >= total)
                break;
            fetch(min(c + 3, last), b0);
            step(c + 1, b1);
            if (c + 2 >= total)
                break;
            fetch(min(c + 4, last), b1);
            step(c + 2, b2);
        }
    }
    else if (DEPTH == 0) {
        for (int c = 0; c < total; c++) {
            fetch(c, b0);
            step(c, b0);
        }
    }
    else {
        fetch(0, b0);
        for (int c = 0; c < total; c += 2) {
            fetch(min(c + 1, last), b1);
            step(c, b0);
            if (c + 1 >= total)
                break;
            fetch(min(c + 2, last), b0);
            step(c + 1, b1);
        }
    }

    /* Pair coherences |T_ij| of the estimated pixels, from LDS. */
    if (kp.npairs) {
#pragma unroll
        for (int qq = 0; qq < PPW; qq++) {
            if (nch[qq] == 0)
                continue;
            for (int t = lane; t < kp.npairs; t += WAVE) {
                const float2 v = tl[qq * NT + (kp.pairs_all ? t : upper_index(N, t, t + 1))];

                pairs[(size_t)t * plane + p0 + qq] = sqrt_fast(v.x * v.x + v.y * v.y);
            }
        }
    }
}

#endif
