// Host-side polyMesh preprocessing: geometry, renumbering, addressing.
//
// Geometry restates the OpenFOAM definitions ICSFoam consumes through
// fvMesh (Sf, magSf, Cf, C, V, weights, deltaCoeffs, nonOrthDeltaCoeffs,
// nonOrthCorrectionVectors); call sites: setCoAndDeltaT.H:61-70,
// convectiveFluxScheme.C:376-417, blockFvOperatorsTemplates.C:499-557.
#include "icsf_mesh.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <numeric>

namespace icsf {

namespace {
inline double mag3(const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
const double ROOTVSMALL = 1.0e-150;
const double VSMALL = 1.0e-300;
}  // namespace

int HostMesh::patchOf(int face) const {
    for (size_t p = 0; p < patchStart.size(); ++p)
        if (face >= patchStart[p] && face < patchStart[p] + patchSize[p]) return (int)p;
    return -1;
}

std::string buildHostMesh(HostMesh& m, int nPoints, const double* points, int nFaces,
                          const int* faceOff, const int* faceVerts, int nInternal,
                          const int* owner, const int* neighbour, int nPatches,
                          const int* patchStart, const int* patchSize, const int* patchKind,
                          const int* patchNbr, const double* patchT) {
    if (nPoints <= 0 || nFaces <= 0 || nInternal < 0 || nInternal > nFaces)
        return "invalid mesh sizes";
    m.nPoints = nPoints;
    m.nFaces = nFaces;
    m.nInternal = nInternal;
    m.points.assign(points, points + 3 * (size_t)nPoints);
    m.faceOff.assign(faceOff, faceOff + nFaces + 1);
    m.faceVerts.assign(faceVerts, faceVerts + faceOff[nFaces]);
    m.owner.assign(owner, owner + nFaces);
    m.neighbour.assign(neighbour, neighbour + nInternal);
    m.patchStart.assign(patchStart, patchStart + nPatches);
    m.patchSize.assign(patchSize, patchSize + nPatches);
    m.patchKind.assign(patchKind, patchKind + nPatches);
    m.patchNbr.assign(nPatches, -1);
    m.patchT.assign(9 * (size_t)nPatches, 0.0);
    for (int p = 0; p < nPatches; ++p) {
        if (patchKind[p] != PATCH_CYCLIC) continue;
        if (!patchNbr || !patchT) return "cyclic patches need a neighbour patch and a transform";
        m.patchNbr[p] = patchNbr[p];
        for (int e = 0; e < 9; ++e) m.patchT[9 * (size_t)p + e] = patchT[9 * (size_t)p + e];
    }
    int nc = 0;
    for (int f = 0; f < nFaces; ++f) nc = std::max(nc, owner[f] + 1);
    for (int f = 0; f < nInternal; ++f) {
        nc = std::max(nc, neighbour[f] + 1);
        if (owner[f] >= neighbour[f]) return "owner must be < neighbour on internal faces";
        if (f > 0 && (owner[f] < owner[f - 1] ||
                      (owner[f] == owner[f - 1] && neighbour[f] < neighbour[f - 1])))
            return "internal faces must be in upper-triangular order";
    }
    m.nCells = nc;
    for (int f = 0; f < nFaces; ++f) {
        if (faceOff[f + 1] - faceOff[f] < 3) return "face with fewer than 3 points";
        for (int k = faceOff[f]; k < faceOff[f + 1]; ++k)
            if (faceVerts[k] < 0 || faceVerts[k] >= nPoints) return "face point index out of range";
    }
    int expect = nInternal;
    for (int p = 0; p < nPatches; ++p) {
        if (patchStart[p] != expect || patchSize[p] < 0) return "patches must tile the boundary faces in order";
        expect += patchSize[p];
    }
    if (expect != nFaces) return "patches must cover all boundary faces";
    for (int p = 0; p < nPatches; ++p) {
        if (patchKind[p] != PATCH_CYCLIC) continue;
        const int q = m.patchNbr[p];
        if (q < 0 || q >= nPatches || q == p || patchKind[q] != PATCH_CYCLIC || m.patchNbr[q] != p)
            return "cyclic patch " + std::to_string(p) + ": neighbour patch must be a distinct cyclic patch "
                   "naming it back";
        if (patchSize[q] != patchSize[p]) return "cyclic patch " + std::to_string(p) + ": size differs from its "
                                                 "neighbour patch";
        const double* A = &m.patchT[9 * (size_t)p];
        const double* Bq = &m.patchT[9 * (size_t)q];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double aat = 0, ab = 0;
                for (int k = 0; k < 3; ++k) {
                    aat += A[3 * i + k] * A[3 * j + k];
                    ab += A[3 * i + k] * Bq[3 * k + j];
                }
                const double id = i == j ? 1.0 : 0.0;
                if (std::fabs(aat - id) > 1e-12)
                    return "cyclic patch " + std::to_string(p) + ": transform is not a rotation";
                if (std::fabs(ab - id) > 1e-12)
                    return "cyclic patch " + std::to_string(p) + ": transform is not the inverse of the "
                           "neighbour patch's";
            }
    }
    computeGeometry(m);
    // checkMesh's "negative volume": an inverted cell (faces ordered the
    // wrong way round) would make every face flux change sign
    for (int c = 0; c < m.nCells; ++c)
        if (!(m.V[c] > 0)) return "cell " + std::to_string(c) + " has non-positive volume (inverted face order?)";
    std::string err = computeCyclic(m);
    if (!err.empty()) return err;
    buildCellFaces(m);
    return "";
}

// Cyclic pairs: face i of a cyclic patch is coupled to face i of its
// neighbour patch; forwardT of the patch maps the neighbour side's vectors
// into this side's frame.  Weights, deltas and deltaCoeffs follow OpenFOAM
// cyclicFvPatch::makeWeights (w = dn/(d + dn), d = nf & (Cf - C)) and
// cyclicFvPatch::delta ((Cf - C) - forwardT & (Cf' - C')), which
// surfaceInterpolation turns into deltaCoeffs, nonOrthDeltaCoeffs and
// nonOrthCorrectionVectors as on internal faces.
std::string computeCyclic(HostMesh& m) {
    const int nB = m.nFaces - m.nInternal;
    m.cycOf.assign(std::max(nB, 0), -1);
    m.cycFace.clear(); m.cycCell.clear(); m.cycPair.clear(); m.cycR.clear(); m.cycDelta.clear();
    for (size_t p = 0; p < m.patchStart.size(); ++p) {
        if (m.patchKind[p] != PATCH_CYCLIC) continue;
        const int q = m.patchNbr[p];
        const double* R = &m.patchT[9 * p];
        for (int i = 0; i < m.patchSize[p]; ++i) {
            const int f = m.patchStart[p] + i, g = m.patchStart[q] + i;
            const int c = m.owner[f], n = m.owner[g];
            const double* Sf = &m.Sf[3 * (size_t)f];
            const double* Sg = &m.Sf[3 * (size_t)g];
            double nf[3], ng[3], d[3], dn[3], Rdn[3], Rng[3];
            for (int k = 0; k < 3; ++k) {
                nf[k] = Sf[k] / m.magSf[f];
                ng[k] = Sg[k] / m.magSf[g];
                d[k] = m.Cf[3 * (size_t)f + k] - m.C[3 * (size_t)c + k];
                dn[k] = m.Cf[3 * (size_t)g + k] - m.C[3 * (size_t)n + k];
            }
            for (int k = 0; k < 3; ++k) {
                Rdn[k] = R[3 * k] * dn[0] + R[3 * k + 1] * dn[1] + R[3 * k + 2] * dn[2];
                Rng[k] = R[3 * k] * ng[0] + R[3 * k + 1] * ng[1] + R[3 * k + 2] * ng[2];
            }
            // paired faces must face each other after the transform, within
            // the matchTolerance the tutorials' cyclic patches set (0.01:
            // tutorials/VKI-LS89/constant/polyMesh/boundary), relative to the
            // face's area for the areas
            constexpr double tol = 1e-2;
            if (std::fabs(Rng[0] + nf[0]) + std::fabs(Rng[1] + nf[1]) + std::fabs(Rng[2] + nf[2]) > tol ||
                std::fabs(m.magSf[f] - m.magSf[g]) > tol * m.magSf[f])
                return "cyclic patch " + std::to_string(p) + " face " + std::to_string(i) +
                       ": does not match face " + std::to_string(i) + " of its neighbour patch";
            const double di = dot3(nf, d), dni = dot3(ng, dn);
            m.weights[f] = dni / (di + dni);
            double delta[3] = {d[0] - Rdn[0], d[1] - Rdn[1], d[2] - Rdn[2]};
            m.deltaCoeffs[f] = 1.0 / mag3(delta);
            m.nonOrthDeltaCoeffs[f] = 1.0 / std::max(dot3(nf, delta), 0.05 * mag3(delta));
            for (int k = 0; k < 3; ++k) m.corrVecs[3 * (size_t)f + k] = nf[k] - delta[k] * m.nonOrthDeltaCoeffs[f];
            m.cycOf[f - m.nInternal] = (int)m.cycFace.size();
            m.cycFace.push_back(f);
            m.cycCell.push_back(n);
            m.cycPair.push_back(g);
            for (int e = 0; e < 9; ++e) m.cycR.push_back(R[e]);
            for (int k = 0; k < 3; ++k) m.cycDelta.push_back(delta[k]);
        }
    }
    const int nC = m.nCells;
    m.cycRowOff.assign(nC + 1, 0);
    for (int f : m.cycFace) m.cycRowOff[m.owner[f] + 1]++;
    for (int c = 0; c < nC; ++c) m.cycRowOff[c + 1] += m.cycRowOff[c];
    m.cycRowK.assign(m.cycFace.size(), 0);
    std::vector<int> pos(m.cycRowOff.begin(), m.cycRowOff.end() - 1);
    for (size_t k = 0; k < m.cycFace.size(); ++k) m.cycRowK[pos[m.owner[m.cycFace[k]]]++] = (int)k;
    return "";
}

void computeGeometry(HostMesh& m) {
    const int nF = m.nFaces, nC = m.nCells;
    const double* p = m.points.data();
    m.Sf.assign(3 * (size_t)nF, 0.0);
    m.Cf.assign(3 * (size_t)nF, 0.0);
    m.magSf.assign(nF, 0.0);
    // Face centres and areas: triangle fan about the point average
    // (OpenFOAM primitiveMesh::makeFaceCentresAndAreas).
    for (int f = 0; f < nF; ++f) {
        const int* fv = m.faceVerts.data() + m.faceOff[f];
        const int n = m.faceOff[f + 1] - m.faceOff[f];
        double* Sf = &m.Sf[3 * (size_t)f];
        double* Cf = &m.Cf[3 * (size_t)f];
        if (n == 3) {
            const double *a = p + 3 * fv[0], *b = p + 3 * fv[1], *c = p + 3 * fv[2];
            double e1[3], e2[3];
            for (int d = 0; d < 3; ++d) {
                Cf[d] = (1.0 / 3.0) * (a[d] + b[d] + c[d]);
                e1[d] = b[d] - a[d];
                e2[d] = c[d] - a[d];
            }
            cross3(e1, e2, Sf);
            for (int d = 0; d < 3; ++d) Sf[d] *= 0.5;
        } else {
            double fc[3] = {p[3 * fv[0]], p[3 * fv[0] + 1], p[3 * fv[0] + 2]};
            for (int k = 1; k < n; ++k)
                for (int d = 0; d < 3; ++d) fc[d] += p[3 * fv[k] + d];
            for (int d = 0; d < 3; ++d) fc[d] /= n;
            double sumN[3] = {0, 0, 0}, sumAc[3] = {0, 0, 0}, sumA = 0;
            for (int k = 0; k < n; ++k) {
                const double* tp = p + 3 * fv[k];
                const double* np_ = p + 3 * fv[k == n - 1 ? 0 : k + 1];
                double c[3], e1[3], e2[3], nn[3];
                for (int d = 0; d < 3; ++d) {
                    c[d] = tp[d] + np_[d] + fc[d];
                    e1[d] = np_[d] - tp[d];
                    e2[d] = fc[d] - tp[d];
                }
                cross3(e1, e2, nn);
                const double a = mag3(nn);
                for (int d = 0; d < 3; ++d) {
                    sumN[d] += nn[d];
                    sumAc[d] += a * c[d];
                }
                sumA += a;
            }
            if (sumA < ROOTVSMALL) {
                for (int d = 0; d < 3; ++d) { Cf[d] = fc[d]; Sf[d] = 0.0; }
            } else {
                for (int d = 0; d < 3; ++d) {
                    Cf[d] = (1.0 / 3.0) * sumAc[d] / sumA;
                    Sf[d] = 0.5 * sumN[d];
                }
            }
        }
        m.magSf[f] = mag3(Sf);
    }
    // Cell centres and volumes: face pyramids about the face-centre average
    // (OpenFOAM primitiveMesh::makeCellCentresAndVols).
    std::vector<double> cEst(3 * (size_t)nC, 0.0);
    std::vector<int> nCF(nC, 0);
    for (int f = 0; f < nF; ++f) {
        const int o = m.owner[f];
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)o + d] += m.Cf[3 * (size_t)f + d];
        nCF[o] += 1;
    }
    for (int f = 0; f < m.nInternal; ++f) {
        const int n = m.neighbour[f];
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)n + d] += m.Cf[3 * (size_t)f + d];
        nCF[n] += 1;
    }
    for (int c = 0; c < nC; ++c)
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)c + d] /= nCF[c];
    m.C.assign(3 * (size_t)nC, 0.0);
    m.V.assign(nC, 0.0);
    for (int f = 0; f < nF; ++f) {
        const int o = m.owner[f];
        const double* Cf = &m.Cf[3 * (size_t)f];
        const double* ce = &cEst[3 * (size_t)o];
        double d[3] = {Cf[0] - ce[0], Cf[1] - ce[1], Cf[2] - ce[2]};
        const double pyr3Vol = dot3(&m.Sf[3 * (size_t)f], d);
        for (int k = 0; k < 3; ++k) {
            const double pc = (3.0 / 4.0) * Cf[k] + (1.0 / 4.0) * ce[k];
            m.C[3 * (size_t)o + k] += pyr3Vol * pc;
        }
        m.V[o] += pyr3Vol;
    }
    for (int f = 0; f < m.nInternal; ++f) {
        const int n = m.neighbour[f];
        const double* Cf = &m.Cf[3 * (size_t)f];
        const double* ce = &cEst[3 * (size_t)n];
        double d[3] = {ce[0] - Cf[0], ce[1] - Cf[1], ce[2] - Cf[2]};
        const double pyr3Vol = dot3(&m.Sf[3 * (size_t)f], d);
        for (int k = 0; k < 3; ++k) {
            const double pc = (3.0 / 4.0) * Cf[k] + (1.0 / 4.0) * ce[k];
            m.C[3 * (size_t)n + k] += pyr3Vol * pc;
        }
        m.V[n] += pyr3Vol;
    }
    for (int c = 0; c < nC; ++c) {
        if (std::fabs(m.V[c]) > VSMALL) {
            for (int d = 0; d < 3; ++d) m.C[3 * (size_t)c + d] /= m.V[c];
        } else {
            for (int d = 0; d < 3; ++d) m.C[3 * (size_t)c + d] = cEst[3 * (size_t)c + d];
        }
        m.V[c] *= (1.0 / 3.0);
    }
    // Interpolation weights, deltaCoeffs, nonOrthDeltaCoeffs and
    // nonOrthCorrectionVectors (OpenFOAM surfaceInterpolation).
    m.weights.assign(nF, 1.0);
    m.deltaCoeffs.assign(nF, 0.0);
    m.nonOrthDeltaCoeffs.assign(nF, 0.0);
    m.corrVecs.assign(3 * (size_t)nF, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        const double* Sf = &m.Sf[3 * (size_t)f];
        const double* Cf = &m.Cf[3 * (size_t)f];
        const double* Co = &m.C[3 * (size_t)m.owner[f]];
        const double* Cn = &m.C[3 * (size_t)m.neighbour[f]];
        double dOwn[3] = {Cf[0] - Co[0], Cf[1] - Co[1], Cf[2] - Co[2]};
        double dNei[3] = {Cn[0] - Cf[0], Cn[1] - Cf[1], Cn[2] - Cf[2]};
        const double SfdOwn = std::fabs(dot3(Sf, dOwn));
        const double SfdNei = std::fabs(dot3(Sf, dNei));
        m.weights[f] = SfdNei / (SfdOwn + SfdNei);
        double delta[3] = {Cn[0] - Co[0], Cn[1] - Co[1], Cn[2] - Co[2]};
        m.deltaCoeffs[f] = 1.0 / mag3(delta);
        double unitArea[3] = {Sf[0] / m.magSf[f], Sf[1] / m.magSf[f], Sf[2] / m.magSf[f]};
        m.nonOrthDeltaCoeffs[f] = 1.0 / std::max(dot3(unitArea, delta), 0.05 * mag3(delta));
        for (int d = 0; d < 3; ++d)
            m.corrVecs[3 * (size_t)f + d] = unitArea[d] - delta[d] * m.nonOrthDeltaCoeffs[f];
    }
    for (int f = m.nInternal; f < nF; ++f) {
        // non-coupled patch delta: nHat*(nHat & (Cf - C))
        const double* Sf = &m.Sf[3 * (size_t)f];
        const double* Cf = &m.Cf[3 * (size_t)f];
        const double* Co = &m.C[3 * (size_t)m.owner[f]];
        double nHat[3] = {Sf[0] / m.magSf[f], Sf[1] / m.magSf[f], Sf[2] / m.magSf[f]};
        double d[3] = {Cf[0] - Co[0], Cf[1] - Co[1], Cf[2] - Co[2]};
        const double nd = dot3(nHat, d);
        double delta[3] = {nHat[0] * nd, nHat[1] * nd, nHat[2] * nd};
        m.deltaCoeffs[f] = 1.0 / mag3(delta);
        m.nonOrthDeltaCoeffs[f] = 1.0 / std::max(dot3(nHat, delta), 0.05 * mag3(delta));
    }
}

void buildCellFaces(HostMesh& m) {
    const int nC = m.nCells;
    m.faceIsEmpty.assign(m.nFaces - m.nInternal, 0);
    m.twoD = false;
    for (size_t p = 0; p < m.patchStart.size(); ++p) {
        if (m.patchKind[p] == PATCH_EMPTY) {
            m.twoD = m.twoD || m.patchSize[p] > 0;
            for (int f = m.patchStart[p]; f < m.patchStart[p] + m.patchSize[p]; ++f)
                m.faceIsEmpty[f - m.nInternal] = 1;
        }
    }
    m.emptyDir = -1;
    if (m.twoD) {
        // direction of the empty patch normals (largest |Sf| component)
        double acc[3] = {0, 0, 0};
        for (int f = m.nInternal; f < m.nFaces; ++f)
            if (m.faceIsEmpty[f - m.nInternal])
                for (int d = 0; d < 3; ++d) acc[d] += std::fabs(m.Sf[3 * (size_t)f + d]);
        m.emptyDir = (int)(std::max_element(acc, acc + 3) - acc);
    }
    std::vector<int> cnt(nC + 1, 0);
    for (int f = 0; f < m.nInternal; ++f) { cnt[m.owner[f] + 1]++; cnt[m.neighbour[f] + 1]++; }
    for (int f = m.nInternal; f < m.nFaces; ++f)
        if (!m.faceIsEmpty[f - m.nInternal]) cnt[m.owner[f] + 1]++;
    for (int c = 0; c < nC; ++c) cnt[c + 1] += cnt[c];
    m.cellFaceOff = cnt;
    m.cellFaces.assign(cnt[nC], 0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    // faces are visited in increasing index, so each list ends up sorted
    for (int f = 0; f < m.nFaces; ++f) {
        if (f >= m.nInternal && m.faceIsEmpty[f - m.nInternal]) continue;
        m.cellFaces[pos[m.owner[f]]++] = f;
        if (f < m.nInternal) m.cellFaces[pos[m.neighbour[f]]++] = f;
    }
    // neighbour entries of face f were appended after owner entries of
    // lower faces; re-sort each list to guarantee face order
    for (int c = 0; c < nC; ++c) std::sort(m.cellFaces.begin() + cnt[c], m.cellFaces.begin() + cnt[c + 1]);
}

void cellCellsCoupled(int nCells, int nInternal, const int* owner, const int* neighbour, int nCoupled,
                      const int* cOwner, const int* cNbr, std::vector<int>& off, std::vector<int>& adj) {
    off.assign(nCells + 1, 0);
    for (int f = 0; f < nInternal; ++f) { off[owner[f] + 1]++; off[neighbour[f] + 1]++; }
    for (int b = 0; b < nCoupled; ++b) off[cOwner[b] + 1]++;
    for (int c = 0; c < nCells; ++c) off[c + 1] += off[c];
    adj.assign(off[nCells], 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int f = 0; f < nInternal; ++f) {
        adj[pos[owner[f]]++] = neighbour[f];
        adj[pos[neighbour[f]]++] = owner[f];
    }
    for (int b = 0; b < nCoupled; ++b) adj[pos[cOwner[b]]++] = cNbr[b];
}

void cellCells(int nCells, int nInternal, const int* owner, const int* neighbour,
               std::vector<int>& off, std::vector<int>& adj) {
    off.assign(nCells + 1, 0);
    for (int f = 0; f < nInternal; ++f) { off[owner[f] + 1]++; off[neighbour[f] + 1]++; }
    for (int c = 0; c < nCells; ++c) off[c + 1] += off[c];
    adj.assign(off[nCells], 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int f = 0; f < nInternal; ++f) {
        adj[pos[owner[f]]++] = neighbour[f];
        adj[pos[neighbour[f]]++] = owner[f];
    }
}

namespace {
// Cuthill-McKee visiting order (OpenFOAM meshTools::bandCompression):
// breadth-first from the lowest unvisited cell, unvisited neighbours
// appended by increasing connectivity (stable).
void cuthillMcKee(int nCells, const std::vector<int>& off, const std::vector<int>& adj, std::vector<int>& order) {
    std::vector<char> visited(nCells, 0);
    order.clear();
    order.reserve(nCells);
    std::deque<int> next;
    std::vector<int> nbrs, weights, idx;
    for (int start = 0; start < nCells; ++start) {
        if (visited[start]) continue;
        next.push_back(start);
        while (!next.empty()) {
            const int cur = next.front();
            next.pop_front();
            if (visited[cur]) continue;
            visited[cur] = 1;
            order.push_back(cur);
            nbrs.clear();
            weights.clear();
            for (int k = off[cur]; k < off[cur + 1]; ++k) {
                const int nb = adj[k];
                if (!visited[nb]) {
                    nbrs.push_back(nb);
                    weights.push_back(off[nb + 1] - off[nb]);
                }
            }
            idx.resize(nbrs.size());
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return weights[a] < weights[b]; });
            for (int i : idx) next.push_back(nbrs[i]);
        }
    }
}
}  // namespace

void rcmOrder(int nCells, int nInternal, const int* owner, const int* neighbour,
              std::vector<int>& newToOld, int nCoupled, const int* cOwner, const int* cNbr) {
    std::vector<int> off, adj, order;
    cellCellsCoupled(nCells, nInternal, owner, neighbour, nCoupled, cOwner, cNbr, off, adj);
    cuthillMcKee(nCells, off, adj, order);
    newToOld.assign(order.rbegin(), order.rend());
}

int greedyColour(int nCells, int nInternal, const int* owner, const int* neighbour,
                 std::vector<int>& colour) {
    // Cells are coloured in Cuthill-McKee visiting order: each cell after the
    // first of its component has its breadth-first parent already coloured,
    // so a bipartite cell graph (any hexahedral mesh) gets exactly 2 colours.
    std::vector<int> off, adj, order;
    cellCells(nCells, nInternal, owner, neighbour, off, adj);
    cuthillMcKee(nCells, off, adj, order);
    colour.assign(nCells, -1);
    int nColours = 0;
    std::vector<int> mark;
    for (int c : order) {
        mark.assign(nColours + 1, 0);
        for (int k = off[c]; k < off[c + 1]; ++k) {
            const int nb = adj[k];
            if (colour[nb] >= 0) mark[colour[nb]] = 1;
        }
        int col = 0;
        while (mark[col]) ++col;
        colour[c] = col;
        nColours = std::max(nColours, col + 1);
    }
    return nColours;
}

void colourMajorOrder(const std::vector<int>& colour, int nColours, std::vector<int>& newToOld) {
    newToOld.clear();
    newToOld.reserve(colour.size());
    for (int col = 0; col < nColours; ++col)
        for (size_t c = 0; c < colour.size(); ++c)
            if (colour[c] == col) newToOld.push_back((int)c);
}

void renumberMesh(int nCells, int nFaces, int nInternal, const int* faceOff,
                  const int* faceVerts, const int* owner, const int* neighbour,
                  const int* newToOld, std::vector<int>& faceVertsOut,
                  std::vector<int>& ownerOut, std::vector<int>& neighbourOut,
                  std::vector<int>& faceNewToOld) {
    std::vector<int> oldToNew(nCells);
    for (int i = 0; i < nCells; ++i) oldToNew[newToOld[i]] = i;
    std::vector<int> a(nInternal), b(nInternal);
    std::vector<char> flip(nInternal, 0);
    for (int f = 0; f < nInternal; ++f) {
        int o = oldToNew[owner[f]], n = oldToNew[neighbour[f]];
        if (o > n) { std::swap(o, n); flip[f] = 1; }
        a[f] = o;
        b[f] = n;
    }
    // upper-triangular order: bucket by new owner (stable), then order each
    // bucket by new neighbour (stable insertion sort; buckets are small)
    std::vector<int> order(nInternal), start(nCells + 1, 0);
    for (int f = 0; f < nInternal; ++f) start[a[f] + 1]++;
    for (int c = 0; c < nCells; ++c) start[c + 1] += start[c];
    {
        std::vector<int> pos(start.begin(), start.end() - 1);
        for (int f = 0; f < nInternal; ++f) order[pos[a[f]]++] = f;
    }
    for (int c = 0; c < nCells; ++c) {
        for (int i = start[c] + 1; i < start[c + 1]; ++i) {
            const int f = order[i];
            int j = i - 1;
            while (j >= start[c] && b[order[j]] > b[f]) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = f;
        }
    }
    ownerOut.assign(nFaces, 0);
    neighbourOut.assign(nInternal, 0);
    faceNewToOld.assign(nFaces, 0);
    faceVertsOut.clear();
    faceVertsOut.reserve(faceOff[nFaces]);
    std::vector<int> newOff;  // same sizes as old per face, reordered
    for (int nf = 0; nf < nInternal; ++nf) {
        const int f = order[nf];
        ownerOut[nf] = a[f];
        neighbourOut[nf] = b[f];
        faceNewToOld[nf] = flip[f] ? -(f + 1) : f;
        const int s = faceOff[f], e = faceOff[f + 1];
        if (!flip[f]) {
            for (int k = s; k < e; ++k) faceVertsOut.push_back(faceVerts[k]);
        } else {
            // OpenFOAM face::reverseFace: keep the first point, reverse the rest
            faceVertsOut.push_back(faceVerts[s]);
            for (int k = e - 1; k > s; --k) faceVertsOut.push_back(faceVerts[k]);
        }
    }
    for (int f = nInternal; f < nFaces; ++f) {
        ownerOut[f] = oldToNew[owner[f]];
        faceNewToOld[f] = f;
        for (int k = faceOff[f]; k < faceOff[f + 1]; ++k) faceVertsOut.push_back(faceVerts[k]);
    }
}

void sweepLevels(int nCells, int nInternal, const int* owner, const int* neighbour,
                 std::vector<int>& fwdLevel, std::vector<int>& bwdLevel, int& nFwd, int& nBwd) {
    std::vector<int> off, adj;
    cellCells(nCells, nInternal, owner, neighbour, off, adj);
    fwdLevel.assign(nCells, 0);
    bwdLevel.assign(nCells, 0);
    nFwd = nCells ? 1 : 0;
    nBwd = nCells ? 1 : 0;
    for (int c = 0; c < nCells; ++c) {
        int l = 0;
        for (int k = off[c]; k < off[c + 1]; ++k)
            if (adj[k] < c) l = std::max(l, fwdLevel[adj[k]] + 1);
        fwdLevel[c] = l;
        nFwd = std::max(nFwd, l + 1);
    }
    for (int c = nCells - 1; c >= 0; --c) {
        int l = 0;
        for (int k = off[c]; k < off[c + 1]; ++k)
            if (adj[k] > c) l = std::max(l, bwdLevel[adj[k]] + 1);
        bwdLevel[c] = l;
        nBwd = std::max(nBwd, l + 1);
    }
}

void buildSell(int nCells, int nInternal, const int* owner, const int* neighbour, SellLayout& L) {
    L.nCells = nCells;
    L.nSlices = (nCells + SLICE - 1) / SLICE;
    // per-row entry lists (column, face, isUpper)
    std::vector<int> off(nCells + 1, 0);
    for (int f = 0; f < nInternal; ++f) { off[owner[f] + 1]++; off[neighbour[f] + 1]++; }
    for (int c = 0; c < nCells; ++c) off[c + 1] += off[c];
    std::vector<long long> ent(off[nCells]);  // (column << 32) | (face << 1) | isLowerSlot
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int f = 0; f < nInternal; ++f) {
        ent[pos[owner[f]]++] = ((long long)neighbour[f] << 32) | ((long long)f << 1) | 0;
        ent[pos[neighbour[f]]++] = ((long long)owner[f] << 32) | ((long long)f << 1) | 1;
    }
    L.rowLen.assign(nCells, 0);
    L.nLower.assign(nCells, 0);
    L.sliceWidth.assign(L.nSlices, 0);
    for (int c = 0; c < nCells; ++c) {
        std::sort(ent.begin() + off[c], ent.begin() + off[c + 1]);
        L.rowLen[c] = off[c + 1] - off[c];
        L.sliceWidth[c / SLICE] = std::max(L.sliceWidth[c / SLICE], L.rowLen[c]);
    }
    L.slicePtr.assign(L.nSlices + 1, 0);
    for (int s = 0; s < L.nSlices; ++s) L.slicePtr[s + 1] = L.slicePtr[s] + L.sliceWidth[s];
    L.col.assign((size_t)L.slicePtr[L.nSlices] * SLICE, -1);
    L.slotUpper.assign(nInternal, -1);
    L.slotLower.assign(nInternal, -1);
    for (int c = 0; c < nCells; ++c) {
        const int s = c / SLICE, lane = c % SLICE;
        for (int k = 0; k < L.rowLen[c]; ++k) {
            const long long e = ent[off[c] + k];
            const int colc = (int)(e >> 32);
            const int f = (int)((e & 0xffffffffLL) >> 1);
            const int lowerSlot = (int)(e & 1);
            const size_t slot = (size_t)(L.slicePtr[s] + k) * SLICE + lane;
            L.col[slot] = colc;
            if (lowerSlot) L.slotLower[f] = (int)slot; else L.slotUpper[f] = (int)slot;
            if (colc < c) L.nLower[c]++;
        }
    }
}

}  // namespace icsf

namespace icsf {

// Topology of an nx*ny*nz hexahedral block in blockMesh order: internal
// faces upper-triangular (+x, +y, +z neighbour of each cell in cell order),
// then the boundary faces patch by patch; sidePatch[s] gives the patch of
// side s in (-x, +x, -y, +y, -z, +z), faces of a patch ordered side by
// side in that order and by cell index within a side.  Face normals point
// out of the owner.  Point index: i + (nx+1)*(j + (ny+1)*k).
void hexBlockTopology(int nx, int ny, int nz, const int* sidePatch, int nPatches, int* faceVerts, int* owner,
                      int* neighbour, int* patchSize) {
    const long NX = nx + 1, NY = ny + 1;
    auto pid = [&](long i, long j, long k) { return (int)(i + NX * (j + NY * k)); };
    auto cid = [&](long i, long j, long k) { return (int)(i + (long)nx * (j + (long)ny * k)); };
    long f = 0;
    auto quad = [&](int a0, int a1, int a2, int a3) {
        int* q = faceVerts + 4 * f;
        q[0] = a0; q[1] = a1; q[2] = a2; q[3] = a3;
    };
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const int c = cid(i, j, k);
                if (i < nx - 1) {
                    quad(pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i + 1, j + 1, k + 1), pid(i + 1, j, k + 1));
                    owner[f] = c; neighbour[f] = c + 1; ++f;
                }
                if (j < ny - 1) {
                    quad(pid(i, j + 1, k), pid(i, j + 1, k + 1), pid(i + 1, j + 1, k + 1), pid(i + 1, j + 1, k));
                    owner[f] = c; neighbour[f] = c + nx; ++f;
                }
                if (k < nz - 1) {
                    quad(pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1));
                    owner[f] = c; neighbour[f] = c + nx * ny; ++f;
                }
            }
    for (int p = 0; p < nPatches; ++p) {
        long before = f;
        for (int s = 0; s < 6; ++s) {
            if (sidePatch[s] != p) continue;
            for (long k = 0; k < nz; ++k)
                for (long j = 0; j < ny; ++j)
                    for (long i = 0; i < nx; ++i) {
                        bool on = (s == 0 && i == 0) || (s == 1 && i == nx - 1) || (s == 2 && j == 0) ||
                                  (s == 3 && j == ny - 1) || (s == 4 && k == 0) || (s == 5 && k == nz - 1);
                        if (!on) continue;
                        switch (s) {
                            case 0: quad(pid(i, j, k), pid(i, j, k + 1), pid(i, j + 1, k + 1), pid(i, j + 1, k)); break;
                            case 1: quad(pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i + 1, j + 1, k + 1), pid(i + 1, j, k + 1)); break;
                            case 2: quad(pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j, k + 1), pid(i, j, k + 1)); break;
                            case 3: quad(pid(i, j + 1, k), pid(i, j + 1, k + 1), pid(i + 1, j + 1, k + 1), pid(i + 1, j + 1, k)); break;
                            case 4: quad(pid(i, j, k), pid(i, j + 1, k), pid(i + 1, j + 1, k), pid(i + 1, j, k)); break;
                            default: quad(pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)); break;
                        }
                        owner[f] = cid(i, j, k);
                        ++f;
                    }
        }
        patchSize[p] = (int)(f - before);
    }
}

}  // namespace icsf
