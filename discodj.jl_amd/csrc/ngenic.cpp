// N-GenIC white noise on the host (GSL ranlxd1, luxury 202), the seed-to-field map of the reference's
// ngenic_field_gsl (NR/src/ics/ngenic_gsl.jl:21-165): a seed table filled along a symmetric spiral
// by a master generator, one generator per (i, j) column, Hermitian fill of the kz = 0 plane,
// Nyquist planes and DC left at zero.  Columns are independent, so they are filled in parallel.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace dj {

namespace {

struct Ranlxd {
    double xd[12];
    double carry;
    int ir, jr, ir_old, pr;
};

constexpr double kOneBit = 1.0 / 281474976710656.0;  // 2^-48

inline int nxt(int i) { return (i + 1) % 12; }

void ranlxd_seed(Ranlxd &st, uint32_t s) {
    if (s == 0) s = 1;
    uint64_t bits = s;
    int xbit[31];
    for (int k = 0; k < 31; k++) {
        xbit[k] = (int)(bits & 1);
        bits >>= 1;
    }
    int ib = 0, jb = 18;
    for (int k = 0; k < 12; k++) {
        double x = 0;
        for (int m = 0; m < 48; m++) {
            x += x + (double)((xbit[ib] + 1) % 2);
            xbit[ib] = (xbit[ib] + xbit[jb]) % 2;
            ib = (ib + 1) % 31;
            jb = (jb + 1) % 31;
        }
        st.xd[k] = kOneBit * x;
    }
    st.carry = 0;
    st.ir = 11;
    st.jr = 7;
    st.ir_old = 0;
    st.pr = 202;
}

// one subtract-with-borrow step x[ir] = x[jr] - x[ir] - carry (lags 12 and 5), advancing ir, jr
inline void swb(Ranlxd &st, int &ir, int &jr, double &carry) {
    double y = st.xd[jr] - st.xd[ir] - carry;
    if (y < 0) { carry = kOneBit; y += 1; } else { carry = 0; }
    st.xd[ir] = y;
    ir = nxt(ir);
    jr = nxt(jr);
}

// skip pr steps of the recurrence after each 12 delivered numbers (luxury level)
void ranlxd_advance(Ranlxd &st) {
    int ir = st.ir, jr = st.jr, k = 0;
    double carry = st.carry;
    for (; ir > 0; k++) swb(st, ir, jr, carry);
    // ir == 0 here: whole blocks of 12 steps, then the remainder
    for (; k + 12 <= st.pr; k += 12)
        for (int t = 0; t < 12; t++) swb(st, ir, jr, carry);
    for (; k < st.pr; k++) swb(st, ir, jr, carry);
    st.ir = ir;
    st.ir_old = ir;
    st.jr = jr;
    st.carry = carry;
}

inline double ranlxd_next(Ranlxd &st) {
    st.ir = nxt(st.ir);
    if (st.ir == st.ir_old) ranlxd_advance(st);
    return st.xd[st.ir];
}

}  // namespace

void ranlxd1_uniforms(uint32_t seed, int n, double *out) {
    Ranlxd st;
    ranlxd_seed(st, seed);
    for (int i = 0; i < n; i++) out[i] = ranlxd_next(st);
}

// The seed of every (i, j) column's generator, tab[i res + j] (ngenic_gsl.jl:103-119): res^2 draws of
// the master generator along the symmetric spiral; sequential, a few milliseconds at res = 1024
void ngenic_seed_table(uint32_t seed, int res, uint32_t *tab_out) {
    const int half = res / 2;
    const size_t R = (size_t)res;
    uint32_t *tab = tab_out;
    std::memset(tab, 0, sizeof(uint32_t) * R * R);
    Ranlxd master;
    ranlxd_seed(master, seed);
    auto draw = [&]() { return (uint32_t)std::floor(2147483647.0 * ranlxd_next(master)); };
    for (int i = 0; i < half; i++) {
        for (int j = 0; j < i; j++) tab[i * R + j] = draw();
        for (int j = 0; j <= i; j++) tab[j * R + i] = draw();
        for (int j = 0; j < i; j++) tab[(R - 1 - i) * R + j] = draw();
        for (int j = 0; j <= i; j++) tab[(R - 1 - j) * R + i] = draw();
        for (int j = 0; j < i; j++) tab[i * R + (R - 1 - j)] = draw();
        for (int j = 0; j <= i; j++) tab[j * R + (R - 1 - i)] = draw();
        for (int j = 0; j < i; j++) tab[(R - 1 - i) * R + (R - 1 - j)] = draw();
        for (int j = 0; j <= i; j++) tab[(R - 1 - j) * R + (R - 1 - i)] = draw();
    }
}

// field (res, res, res/2+1) complex interleaved, element (i, j, k) at i + res (j + res k)
void ngenic_field(uint32_t seed, int res, double *out) {
    const int half = res / 2, nh = res / 2 + 1;
    const size_t R = (size_t)res;
    std::memset(out, 0, sizeof(double) * 2 * R * R * nh);
    std::vector<uint32_t> tab(R * R, 0);
    ngenic_seed_table(seed, res, tab.data());
    const double twopi = 2 * M_PI;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int i = 0; i < res; i++)
        for (int j = 0; j < res; j++) {
            const int ii = i > 0 ? res - i : 0, jj = j > 0 ? res - j : 0;
            Ranlxd st;
            ranlxd_seed(st, tab[i * R + j]);
            for (int k = 0; k <= half; k++) {
                double phase = ranlxd_next(st) * twopi, ampl;
                do { ampl = ranlxd_next(st); } while (ampl == 0.0 || ampl == 1.0);
                if (i == half || j == half || k == half) continue;
                if (i == 0 && j == 0 && k == 0) continue;
                ampl = std::sqrt(-std::log(ampl));
                double re = ampl * std::cos(phase), im = ampl * std::sin(phase);
                auto put = [&](int a, int b, int c, double x, double y) {
                    size_t e = (size_t)a + R * ((size_t)b + R * c);
                    out[2 * e] = x;
                    out[2 * e + 1] = y;
                };
                if (k > 0) {
                    put(i, j, k, re, im);
                } else if (i == 0) {
                    if (j < half) {
                        put(0, j, 0, re, im);
                        put(0, jj, 0, re, -im);
                    }
                } else if (i < half) {
                    put(i, j, 0, re, im);
                    put(ii, jj, 0, re, -im);
                }
            }
        }
}

}  // namespace dj
