// Harmonic-balance time-spectral operator (host side).
//   omega list      HBZone::setOmegaList        (HBZone.C:46-91)
//   instants        HBZoneList::setInstants     (HBZoneList.C:253-346)
//   condition no.   HBZone::calcConditionN      (HBZone.C:195-268)
//   D               HBZone::updateHBOperators   (HBZone.C:270-356)
#pragma once
#include <string>
#include <vector>

namespace icsf {

// 0, then +k f_i (k = 1..h_i, frequencies in order), then the negative
// frequencies in reverse order.  Frequencies are angular (rad/s); their
// sign is dropped.
std::vector<double> hbOmegaList(const std::vector<double>& freqs, const std::vector<int>& harmonics);

// Period of the lowest non-zero angular frequency, 2 pi / min |omega|.
double hbT0(const std::vector<double>& omegas);

// L2 condition number of E^-1 = [exp(i omega_k t_n)] as the reference
// evaluates it: square roots of the eigenvalues of Re((E^-1)^H E^-1)
// (Matrix::T() is the conjugate transpose), GREAT when the smallest is
// below SMALL.
double hbConditionNumber(const std::vector<double>& omegas, const std::vector<double>& t);

// Equispaced instants t_n = n T/N over the selected period, or (T <= 0)
// the period in [T0, 5 T0] (steps of 0.001 T0) with the lowest condition
// number.  Returns the instants.
std::vector<double> hbInstants(int nInstants, double selectedPeriod, const std::vector<double>& omegas,
                               double* condOut = nullptr);

// D[nO][nO] = -(Im(E^-1) A Re(E) + Re(E^-1) A Im(E)), E = pinv(E^-1),
// A = diag(omega).  Returns false (with a message) if E^-1 is rank
// deficient.
bool hbOperator(const std::vector<double>& omegas, const std::vector<double>& t, std::vector<double>& D,
                std::string& err);

// Time shift of a phase-lagged cyclic patch side
// (phaseLagCyclicFvPatchField.C:289-326): Dl[nO][nO] = Re(E^-1 M E), E =
// pinv(E^-1), M = diag(1, e^{i n IBPA} (n = 1..nH), e^{-i n IBPA} at
// nT - n) over the omega list's indices.  A field sampled at the instants
// is moved by IBPA / omega_1 in time: sum_m Dl[l][m] x(t_m) = x(t_l +
// IBPA/omega_1) for x in the span of the omega list.
bool hbLagOperator(const std::vector<double>& omegas, const std::vector<double>& t, double ibpa,
                   std::vector<double>& Dl, std::string& err);

// Weights of the Fourier reconstruction at `time` (HBZone::reconstruct,
// HBZoneTemplates.C:328-365): w[m] = Re(sum_n exp(i omega_n time) E[n][m]),
// E = pinv(E^-1); a field at that time is sum_m w[m] field_m.
bool hbReconstructWeights(const std::vector<double>& omegas, const std::vector<double>& t, double time,
                          std::vector<double>& w, std::string& err);

}  // namespace icsf
