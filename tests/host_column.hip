// The device column routines (xcape_amd/csrc/cape_column.h) compiled for the
// host, so that the CPU test suite can check them against the oracle without
// a GPU.  Host math (glibc expf/logf/powf) is what the oracle uses too, so the
// plain method must agree bit for bit.
#include <math.h>

#include "../xcape_amd/csrc/cape_column.h"

using namespace xcape;

namespace {

// level k of a column, k = 0 the surface; `ls` elements between levels
struct HostCol {
    const float *p, *t, *td;
    float ps, ts, tds;
    long ls, pls;
    float P(int k) const { return k == 0 ? ps : p[(k - 1) * pls]; }
    float Tc(int k) const { return k == 0 ? ts : t[(k - 1) * ls]; }
    float Tdc(int k) const { return k == 0 ? tds : td[(k - 1) * ls]; }
    Raw raw_above(int k) const { return Raw{p[(k - 1) * pls], t[(k - 1) * ls], td[(k - 1) * ls]}; }
};

template <bool ICE, bool PSEUDO>
void run(const HostCol &c, int nk, float pinc, int source, float ml_depth, int warm,
         ColumnOut<float> &o, LaneStats &st)
{
    if (warm) {
        Parcel<float> pc;
        if (parcel_source<float, MATH_ACCURATE>(c, nk, source, ml_depth, pc))
            cape_column_warm<float, MATH_ACCURATE, ICE, PSEUDO, true>(c, nk, pinc, pc, NAN,
                                                                      1e-3f, o, st);
        else {
            o.cape = 0.0f; o.cin = 0.0f; o.mulvl = pc.mulvl; o.zout = -999999.0f;
            o.p_lfc = NAN; o.p_el = NAN;
        }
    } else {
        cape_column<float, MATH_ACCURATE, ICE, PSEUDO, HostCol, NoLog, true>(
            c, nk, pinc, source, ml_depth, o, st);
    }
}

}  // namespace

// p1d != NULL: shared pressure levels with 1-based start levels; level_major:
// (nk, ncol) profile arrays.  lfc, el: the LFC/EL pressures (hPa), or NULL.
// stats: levels, substeps, iterations, nonconverged, fallbacks,
// first-sub-step iterations.
extern "C" void host_cape(const float *p, const float *p1d, const int *start, const float *t,
                          const float *td, const float *ps, const float *ts, const float *tds,
                          int nk_in, int ncol, int level_major, float pinc, int source,
                          float ml_depth, int adiabat, int warm, float *cape, float *cin,
                          int *mul, float *z, float *lfc, float *el, long long *stats)
{
    LaneStats st;
    for (int i = 0; i < ncol; ++i) {
        const long ls = level_major ? ncol : 1;
        long off = level_major ? i : (long)i * nk_in;
        HostCol c{};
        int nk = nk_in + 1;
        if (p1d) {
            const int s0 = start[i] - 1;
            c.p = p1d + s0;
            c.pls = 1;
            off += s0 * ls;
            nk = nk_in - s0 + 1;
        } else {
            c.p = p + off;
            c.pls = ls;
        }
        c.t = t + off;
        c.td = td + off;
        c.ls = ls;
        c.ps = ps[i]; c.ts = ts[i]; c.tds = tds[i];
        ColumnOut<float> o;
        if (!(ts[i] > 0.0f)) {
            cape[i] = cin[i] = z[i] = 0.0f;
            mul[i] = 0;
            if (lfc) { lfc[i] = NAN; el[i] = NAN; }
            continue;
        }
        switch (adiabat) {
        case 1: run<false, true>(c, nk, pinc, source, ml_depth, warm, o, st); break;
        case 2: run<false, false>(c, nk, pinc, source, ml_depth, warm, o, st); break;
        case 3: run<true, true>(c, nk, pinc, source, ml_depth, warm, o, st); break;
        default: run<true, false>(c, nk, pinc, source, ml_depth, warm, o, st); break;
        }
        cape[i] = o.cape; cin[i] = o.cin; mul[i] = o.mulvl; z[i] = o.zout;
        if (lfc) { lfc[i] = o.p_lfc; el[i] = o.p_el; }
    }
    stats[0] = st.levels; stats[1] = st.substeps; stats[2] = st.iterations;
    stats[3] = st.nonconverged; stats[4] = st.fallbacks; stats[5] = st.first;
}
