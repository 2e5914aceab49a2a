// Harmonic-balance time-spectral operator, FP64 on the host (built once per
// case; N <= a few tens).  Complex arithmetic with std::complex; the
// pseudo-inverse is the exact inverse (Gauss-Jordan, partial pivoting) for
// square E^-1 and (E^H E)^-1 E^H for the oversampled (tall) case; the
// symmetric eigenvalues of the condition-number test come from cyclic
// Jacobi rotations.
#include "icsf_hb.h"

#include <algorithm>
#include <cmath>
#include <complex>

namespace icsf {

namespace {

typedef std::complex<double> cplx;
constexpr double kSmall = 1.0e-15;
constexpr double kGreat = 1.0e15;
constexpr double kPi = 3.14159265358979323846;

// E^-1[n][k] = exp(i omega_k t_n), row-major [nO][nT]
std::vector<cplx> synthesis(const std::vector<double>& omegas, const std::vector<double>& t) {
    const size_t nO = t.size(), nT = omegas.size();
    std::vector<cplx> E(nO * nT);
    for (size_t n = 0; n < nO; ++n)
        for (size_t k = 0; k < nT; ++k) E[n * nT + k] = cplx(std::cos(omegas[k] * t[n]), std::sin(omegas[k] * t[n]));
    return E;
}

// In-place inverse of a square complex matrix (row-major n x n).
bool invert(std::vector<cplx>& A, size_t n) {
    std::vector<cplx> I(n * n, cplx(0, 0));
    for (size_t i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (size_t col = 0; col < n; ++col) {
        size_t piv = col;
        double best = std::abs(A[col * n + col]);
        for (size_t r = col + 1; r < n; ++r)
            if (std::abs(A[r * n + col]) > best) { best = std::abs(A[r * n + col]); piv = r; }
        if (best < 1e-13) return false;
        if (piv != col)
            for (size_t k = 0; k < n; ++k) { std::swap(A[piv * n + k], A[col * n + k]); std::swap(I[piv * n + k], I[col * n + k]); }
        const cplx d = A[col * n + col];
        for (size_t k = 0; k < n; ++k) { A[col * n + k] /= d; I[col * n + k] /= d; }
        for (size_t r = 0; r < n; ++r) {
            if (r == col) continue;
            const cplx f = A[r * n + col];
            if (f == cplx(0, 0)) continue;
            for (size_t k = 0; k < n; ++k) { A[r * n + k] -= f * A[col * n + k]; I[r * n + k] -= f * I[col * n + k]; }
        }
    }
    A.swap(I);
    return true;
}

// Eigenvalues of a real symmetric matrix (cyclic Jacobi).
std::vector<double> symEig(std::vector<double> A, size_t n) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (size_t i = 0; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        if (off < 1e-30) break;
        for (size_t p = 0; p < n; ++p)
            for (size_t q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                const double tt = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(tt * tt + 1), s = tt * c;
                for (size_t k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (size_t k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    std::vector<double> ev(n);
    for (size_t i = 0; i < n; ++i) ev[i] = A[i * n + i];
    return ev;
}

}  // namespace

std::vector<double> hbOmegaList(const std::vector<double>& freqs, const std::vector<int>& harmonics) {
    std::vector<double> om{0.0}, pos, neg;
    for (size_t i = 0; i < freqs.size(); ++i) {
        const double f = std::fabs(freqs[i]);
        for (int k = 1; k <= harmonics[i]; ++k) {
            pos.push_back(k * f);
            neg.push_back(-k * f);
        }
    }
    om.insert(om.end(), pos.begin(), pos.end());
    om.insert(om.end(), neg.rbegin(), neg.rend());
    return om;
}

double hbT0(const std::vector<double>& omegas) {
    double mn = kGreat;
    for (double w : omegas)
        if (std::fabs(w) > 0 && std::fabs(w) < mn) mn = std::fabs(w);
    return 2 * kPi / mn;
}

double hbConditionNumber(const std::vector<double>& omegas, const std::vector<double>& t) {
    const size_t nO = t.size(), nT = omegas.size();
    const std::vector<cplx> E = synthesis(omegas, t);
    std::vector<double> M(nT * nT, 0.0);
    for (size_t l = 0; l < nT; ++l)
        for (size_t m = 0; m < nT; ++m) {
            cplx s(0, 0);
            for (size_t n = 0; n < nO; ++n) s += std::conj(E[n * nT + l]) * E[n * nT + m];
            M[l * nT + m] = s.real();
        }
    const std::vector<double> ev = symEig(M, nT);
    const double mn = *std::min_element(ev.begin(), ev.end());
    const double mx = *std::max_element(ev.begin(), ev.end());
    if (mn < kSmall) return kGreat;
    return std::sqrt(mx) / std::sqrt(mn);
}

std::vector<double> hbInstants(int nInstants, double selectedPeriod, const std::vector<double>& omegas,
                               double* condOut) {
    std::vector<double> t(nInstants, 0.0);
    if (selectedPeriod > 0) {
        const double dT = selectedPeriod / nInstants;
        for (int n = 0; n < nInstants; ++n) t[n] = n * dT;
        if (condOut) *condOut = hbConditionNumber(omegas, t);
        return t;
    }
    const double T0 = hbT0(omegas), Tf = 5 * T0, step = 0.001 * T0;
    std::vector<double> periods{T0};
    double Ti = T0;
    while (Ti <= Tf) {
        Ti += step;
        periods.push_back(Ti);
    }
    double best = kGreat;
    std::vector<double> snap(nInstants), sel(nInstants, 0.0);
    for (double P : periods) {
        const double dT = P / nInstants;
        for (int l = 0; l < nInstants; ++l) snap[l] = dT * l;
        double c = hbConditionNumber(omegas, snap);
        if (c < kSmall) c = kSmall;
        if (c < best) { best = c; sel = snap; }
    }
    if (condOut) *condOut = best;
    return sel;
}

namespace {
// E = pinv(E^-1) [nT][nO] (HBZone.C:292): the inverse when square, else
// ((E^-1)^H E^-1)^-1 (E^-1)^H
bool hbPinv(const std::vector<double>& omegas, const std::vector<double>& t, const std::vector<cplx>& Einv,
            std::vector<cplx>& E, std::string& err) {
    const size_t nO = t.size(), nT = omegas.size();
    if (nO < nT) { err = "HB: fewer instants than frequencies"; return false; }
    E.assign(nT * nO, cplx(0, 0));
    if (nO == nT) {
        std::vector<cplx> A = Einv;
        if (!invert(A, nT)) { err = "HB: E^-1 is singular for these instants"; return false; }
        E = A;
    } else {
        std::vector<cplx> N(nT * nT, cplx(0, 0));  // (E^-1)^H E^-1
        for (size_t l = 0; l < nT; ++l)
            for (size_t m = 0; m < nT; ++m)
                for (size_t n = 0; n < nO; ++n) N[l * nT + m] += std::conj(Einv[n * nT + l]) * Einv[n * nT + m];
        if (!invert(N, nT)) { err = "HB: E^-1 is rank deficient for these instants"; return false; }
        for (size_t l = 0; l < nT; ++l)
            for (size_t n = 0; n < nO; ++n) {
                cplx s(0, 0);
                for (size_t m = 0; m < nT; ++m) s += N[l * nT + m] * std::conj(Einv[n * nT + m]);
                E[l * nO + n] = s;
            }
    }
    return true;
}
}  // namespace

bool hbOperator(const std::vector<double>& omegas, const std::vector<double>& t, std::vector<double>& D,
                std::string& err) {
    const size_t nO = t.size(), nT = omegas.size();
    const std::vector<cplx> Einv = synthesis(omegas, t);  // [nO][nT]
    std::vector<cplx> E;                                   // pinv, [nT][nO]
    if (!hbPinv(omegas, t, Einv, E, err)) return false;
    // temp_0 = Im(E^-1) A, temp_1 = Re(E^-1) A ; d_0 = temp_0 Re(E), d_1 = temp_1 Im(E)
    std::vector<double> t0(nO * nT, 0.0), t1(nO * nT, 0.0), d0(nO * nO, 0.0), d1(nO * nO, 0.0);
    for (size_t l = 0; l < nO; ++l)
        for (size_t m = 0; m < nT; ++m)
            for (size_t n = 0; n < nT; ++n) {
                const double A = (n == m) ? omegas[n] : 0.0;
                t0[l * nT + m] += Einv[l * nT + n].imag() * A;
                t1[l * nT + m] += Einv[l * nT + n].real() * A;
            }
    for (size_t l = 0; l < nO; ++l)
        for (size_t m = 0; m < nO; ++m)
            for (size_t n = 0; n < nT; ++n) {
                d0[l * nO + m] += t0[l * nT + n] * E[n * nO + m].real();
                d1[l * nO + m] += t1[l * nT + n] * E[n * nO + m].imag();
            }
    D.assign(nO * nO, 0.0);
    for (size_t i = 0; i < nO * nO; ++i) D[i] = -(d0[i] + d1[i]);
    return true;
}

bool hbLagOperator(const std::vector<double>& omegas, const std::vector<double>& t, double ibpa,
                   std::vector<double>& Dl, std::string& err) {
    const size_t nO = t.size(), nT = omegas.size();
    const std::vector<cplx> Einv = synthesis(omegas, t);  // [nO][nT]
    std::vector<cplx> E;                                   // [nT][nO]
    if (!hbPinv(omegas, t, Einv, E, err)) return false;
    // M = diag(1, e^{i n IBPA} (n = 1..nH), e^{-i n IBPA} at nT - n): the
    // product M E scales row l of E by M[l][l] (the off-diagonal zeros add
    // nothing to the sums)
    const size_t nH = (nT - 1) / 2;
    std::vector<cplx> mdiag(nT, cplx(0, 0));
    mdiag[0] = cplx(1, 0);
    for (size_t n = 1; n <= nH; ++n) {
        const cplx e(std::cos((double)n * ibpa), std::sin((double)n * ibpa));
        mdiag[n] = e;
        mdiag[nT - n] = std::conj(e);
    }
    Dl.assign(nO * nO, 0.0);
    for (size_t l = 0; l < nO; ++l)
        for (size_t m = 0; m < nO; ++m) {
            cplx d(0, 0);
            for (size_t n = 0; n < nT; ++n) d += Einv[l * nT + n] * (mdiag[n] * E[n * nO + m]);
            Dl[l * nO + m] = d.real();
        }
    return true;
}

bool hbReconstructWeights(const std::vector<double>& omegas, const std::vector<double>& t, double time,
                          std::vector<double>& w, std::string& err) {
    const size_t nO = t.size(), nT = omegas.size();
    std::vector<cplx> E;
    if (!hbPinv(omegas, t, synthesis(omegas, t), E, err)) return false;
    std::vector<cplx> e1(nT);
    for (size_t k = 0; k < nT; ++k) e1[k] = cplx(std::cos(omegas[k] * time), std::sin(omegas[k] * time));
    w.assign(nO, 0.0);
    for (size_t m = 0; m < nO; ++m) {
        cplx s(0, 0);
        for (size_t n = 0; n < nT; ++n) s += e1[n] * E[n * nO + m];
        w[m] = s.real();
    }
    return true;
}

}  // namespace icsf
