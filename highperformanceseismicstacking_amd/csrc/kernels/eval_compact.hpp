// OCT evaluation with compacted trace passes, for the differential-evolution kernel.
//
// In an OCT evaluation most traces fall outside the aperture |m - mh| <= apm around the
// candidate's displaced midpoint mh (about three in four on a line whose selected traces
// span the reachable displacements).  With one candidate per lane, every lane of a wave
// has its own mh, so in evaluate_n's trace loop some lane keeps nearly every trace and
// the whole wave pays the window read for it.  Here the loop over a run of traces that
// share a half offset (the lists are grouped by offset) goes in two passes:
//   1. the aperture test of every trace of the run for every candidate, collected into a
//      per-lane bit mask (and the rejections counted as evaluate_n counts them);
//   2. each lane walks the set bits of its own mask in increasing trace order, taking the
//      trace's midpoint, scale and sample offset from the wave's LDS record of the chunk,
//      so the wave runs as many window reads as its busiest lane keeps, not as the union
//      of all lanes keeps.
// A lane adds its kept traces in the same order as evaluate_n and computes every value
// with the same operations, so the results are the same bits (HSS_EXACT and HSS_FAST).
// HSS_KAHAN flushes its chunk sums at fixed trace-block boundaries of the list, which a
// per-lane walk does not visit, so that variant stays on evaluate_n.
#pragma once

#include "eval.hpp"

namespace hss {

// One trace of the current chunk as the second pass needs it (lane j writes trace j0+j).
struct alignas(16) OctTraceRec {
    float m;
    float sc;
    uint32_t off_lo, off_hi;  // first sample of the trace in the flat array
};

template <int W, int VARIANT, int STORAGE, bool SHIFT, int CPL>
__device__ __forceinline__ void evaluate_oct_compact(const EvalCtx &e, float t0,
                                                     const Candidate (&k)[CPL],
                                                     const bool (&on)[CPL], float shift,
                                                     bool active, OctTraceRec *recs,
                                                     float (&S)[CPL], float (&stack)[CPL],
                                                     unsigned *rejected) {
    static_assert(VARIANT != HSS_KAHAN, "the compensated variant uses evaluate_n");
    using Reader = SampleReader<STORAGE>;
    constexpr bool FASTIDX = W != 0;
    const bool scaled = (STORAGE != HSS_F32) && e.scale != nullptr;
    Reader rd{(decltype(Reader::p))e.samples};
    Window<W, VARIANT> acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].reset();
    unsigned rej = 0;  // the wave's rejections (wave-uniform)
    const int lane = threadIdx.x & (kWave - 1);
    OctCand oc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) oc[c] = oct_cand(t0, e.h0, k[c]);

    for (int j0 = 0; j0 < e.cnt; j0 += kWave) {
        const int mine = j0 + lane;
        const bool have = mine < e.cnt;
        const int ti_l = have ? (e.midx ? e.midx[e.beg + mine] : e.beg + mine) : 0;
        const float m_l = have ? e.mid[ti_l] : 0.0f;
        const float h_l = have ? e.hoff[ti_l] : 0.0f;
        const float sc_l = (scaled && have) ? e.scale[ti_l] : 1.0f;
        const OctTrace o_l = oct_trace<FASTIDX>(h_l, m_l, e.h0);
        const size_t off_l = (size_t)ti_l * (size_t)e.ns;
        wave_lds_order();  // the previous chunk's records have been read
        recs[lane] = OctTraceRec{m_l, sc_l, (uint32_t)off_l, (uint32_t)(off_l >> 32)};
        wave_lds_order();
        const int nchunk = min(kWave, e.cnt - j0);
        const uint32_t hkey_l = __float_as_uint(h_l);
        int jj = 0;
        while (jj < nchunk) {
            // the run jj .. je-1 of traces with trace jj's half offset (its bit pattern)
            const uint32_t key = (uint32_t)__builtin_amdgcn_readlane((int)hkey_l, jj);
            const unsigned long long later = ~0ull << jj << 1;  // lanes above jj (jj < 64)
            const unsigned long long differ = __builtin_amdgcn_ballot_w64(hkey_l != key) & later;
            const int je = differ ? min(__builtin_ctzll(differ), nchunk) : nchunk;
            OctTrace ot{};
            ot.h = bcast(o_l.h, jj);
            bcast_oct_mid_terms(ot, o_l, jj);
            OctMid om[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) om[c] = oct_mid<FASTIDX>(ot, oc[c], e.m0);
            // pass 1: aperture decisions, as oct_traveltime_at takes them
            uint32_t klo[CPL], khi[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) klo[c] = khi[c] = 0u;
            for (int t = jj; t < je; ++t) {
                const float m = bcast(m_l, t);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const bool r = om[c].bad | (fabsf(m - om[c].mh) > e.apm);
                    rej += (unsigned)__builtin_popcountll(
                        __builtin_amdgcn_ballot_w64(r & on[c] & active));
                    const bool keep = !r & on[c] & active;
                    if (t < 32)
                        klo[c] |= keep ? (1u << t) : 0u;
                    else
                        khi[c] |= keep ? (1u << (t - 32)) : 0u;
                }
            }
            // the run's time terms (from trace jj, which has the run's offset), where some
            // lane keeps a trace for the candidate
            OctTime otm[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                otm[c] = OctTime{0.0f, 0.0f, true};
                if (__builtin_amdgcn_ballot_w64((klo[c] | khi[c]) != 0u) != 0) {
                    bcast_oct_time_terms(ot, o_l, jj);
                    otm[c] = oct_time_terms<FASTIDX>(ot, oc[c], e.m0, om[c]);
                }
                if (otm[c].bad) klo[c] = khi[c] = 0u;
            }
            // pass 2: each lane's kept traces in increasing order (the next one is the
            // lowest set bit; selections instead of branches, one masked window read per
            // candidate and step)
            // steps: the largest number of kept traces of any lane (binary lifting on
            // ballots), so the walk is a counted, wave-uniform loop
            int most = 0;
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                most = max(most, __builtin_popcount(klo[c]) + __builtin_popcount(khi[c]));
            int steps = 0;
#pragma unroll
            for (int b = 64; b > 0; b >>= 1)
                if (__builtin_amdgcn_ballot_w64(most >= steps + b) != 0) steps += b;
            for (int step = 0; step < steps; ++step) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const bool has = (klo[c] | khi[c]) != 0u;
                    const bool low = klo[c] != 0u;
                    const uint32_t word = low ? klo[c] : khi[c];
                    const int t = ((int)__builtin_ctz(word | 0x80000000u) + (low ? 0 : 32)) & 63;
                    const uint32_t rest = word & (word - 1u);
                    klo[c] = low ? rest : klo[c];
                    khi[c] = low ? khi[c] : rest;
                    const OctTraceRec rec = recs[t];
                    float tt = otm[c].th + otm[c].ah * (rec.m - om[c].mh);
                    if constexpr (SHIFT) tt = tt + shift;
                    int kk;
                    float f;
                    const bool in = window_k<FASTIDX>(tt, e.dt, e.inv_dt, e.ib, &kk, &f);
                    if (has & in) {
                        const size_t trace0 = ((size_t)rec.off_hi << 32) | (size_t)rec.off_lo;
                        if constexpr (STORAGE == HSS_F32) {
                            const float *row = (const float *)e.samples + trace0 - (size_t)e.tau;
                            acc[c].add_row(row, (unsigned)kk, f, e.w);
                        } else if constexpr (W > 0) {
                            const size_t row = trace0 - (size_t)e.tau;
                            const unsigned rpar = (unsigned)row & 1u;
                            const uint32_t *roww =
                                (const uint32_t *)((const uint16_t *)e.samples + (row - rpar));
                            acc[c].template add16<STORAGE>(roww, rpar, (unsigned)kk, f, rec.sc);
                        } else {
                            acc[c].add(rd, trace0 + (size_t)(kk - e.tau), f, rec.sc, scaled, e.w);
                        }
                    }
                }
            }
            jj = je;
        }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].reduce(e.w, &S[c], &stack[c]);
    // the wave's count goes to its lowest active lane
    const unsigned long long act = __builtin_amdgcn_ballot_w64(active);
    *rejected += (act != 0 && lane == __builtin_ctzll(act)) ? rej : 0u;
}

}  // namespace hss
