// SimPLIC alpha-transport ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference's per-time-step alpha transport
// (daidezhi/geometricVofExt, src/SimPLIC), written to follow the reference
// files line by line so that the HIP product path can be checked against it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library; the product never links or calls it.
//
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
// No FMA contraction: every a*b+c is rounded twice, as on x86-64 without -mfma.
//
// Function -> reference citation (paths relative to the reference root):
//   calcSubFace / calcSubFaceCentreAndArea  src/SimPLIC/cut/cutFace/cutFace.C:37-96,136-259
//   timeIntegratedFaceFlux                  src/SimPLIC/cut/cutFace/cutFace.C:262-389
//   timeIntegratedArea                      src/SimPLIC/cut/cutFace/cutFace.C:392-506
//   calcInterfaceCentreAndArea              src/SimPLIC/cut/cutCell/cutCell.C:37-100
//   calcSubCellCentreAndVolume              src/SimPLIC/cut/cutCell/cutCell.C:103-137
//   getLocalPointFieldAndFaceList           src/SimPLIC/cut/cutCell/cutCell.C:140-236
//   calcSubCell                             src/SimPLIC/cut/cutCell/cutCell.C:343-542
//   findSignedDistance                      src/SimPLIC/cut/cutCell/cutCell.C:611-799
//   initialize / reconstruct                src/SimPLIC/reconstruction/reconstruction.C:634-722
//   calcInterfaceNFromIsoAlphaGrad (LS)     src/SimPLIC/reconstruction/reconstruction.C:85-141
//   updateFaceFlatness                      src/SimPLIC/reconstruction/reconstruction.C:408-440
//   timeIntegratedFlux                      src/SimPLIC/advection/advection.C:85-221
//   extendMarkedCells / setDownwindFaces / netFlux / applyBruteForceBounding
//                                           src/SimPLIC/advection/advection.C:54-82,224-308
//   limitFlux / boundFlux / advect          src/SimPLIC/advection/advectionTemplates.C:118-418
//   interfacePoints / interface (PLIC)     src/SimPLIC/cut/cutCell/cutCell.C:545-608,
//                                           src/SimPLIC/reconstruction/reconstruction.C:787-835
//   mapAlphaField                           src/SimPLIC/reconstruction/reconstruction.C:725-784
//   getRhoPhi, sub-cycled alpha step        src/SimPLIC/advection/advection.H:329-343,
//                                           applications/solvers/interPlicFoam/alphaEqnSubCycle.H:30-60
//   harness (U, phi, Courant numbers)       applications/test/plicVofAdvectionFoam/updateU.H:1-70,
//                                           alphaCourantNo.H:87-107
//
// OpenFOAM v2312 pieces the path calls are not in the reference tree; they are
// restated from OpenFOAM's published algorithms (primitiveMeshTools face and
// cell geometry, face::centre/areaNormal, upwind + surfaceInterpolationScheme
// lambda form, surfaceIntegrate, cell-point-cell stencil, polyDegree1 least
// squares with Cholesky inverse, volPointInterpolation inverse-distance point
// values, interpolationCellPoint tet search).  Their bit-level fidelity to
// OpenFOAM is unpinned; each is pinned here by closed forms and invariants in
// tests/test_oracle_*.py.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// Planted mistakes for the pinning tests (tests/test_oracle_pins.py): a test
// switches one on to show that its check fails for a plausible wrong variant.
// 0 = none (the restatement); 1 = least squares weighted by 1/(|d|^2+1e-3);
// 2 = cell-point interpolation in the worst tet instead of the one holding the
// position; 3 = boundFlux walking the cells in descending order.
static int g_mutation = 0;

// OpenFOAM scalar constants (double precision build)
static const double SMALL = 1e-15;
static const double VSMALL = 1e-300;
static const double ROOTVSMALL = 1e-150;
static const double TSMALL = 10.0 * SMALL;   // 1e-14, cutFace.C:154, cutCell.C:355

struct V3 {
    double x, y, z;
};
static inline V3 v3(double a, double b, double c) { return V3{a, b, c}; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
static inline V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
// OpenFOAM inner product a & b
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// OpenFOAM cross product a ^ b
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double magSqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double mag(V3 a) { return std::sqrt(magSqr(a)); }
static inline double sgn(double s) { return (s >= 0.0) ? 1.0 : -1.0; }   // OpenFOAM sign()
static inline double pos0(double s) { return (s >= 0.0) ? 1.0 : 0.0; }
static inline double neg0(double s) { return (s <= 0.0) ? 1.0 : 0.0; }
static const V3 VZERO = {0.0, 0.0, 0.0};

// ---------------------------------------------------------------------------
// OpenFOAM face geometry of an ordered point list (face.C, triangle.H)
// ---------------------------------------------------------------------------

// face::centre
static V3 faceCentre(const std::vector<V3>& p) {
    const int n = (int)p.size();
    if (n == 3) return (1.0 / 3.0) * (p[0] + p[1] + p[2]);
    V3 centrePoint = VZERO;
    for (int i = 0; i < n; ++i) centrePoint += p[i];
    centrePoint = centrePoint / (double)n;
    double sumA = 0.0;
    V3 sumAc = VZERO;
    for (int i = 0; i < n; ++i) {
        const V3& nextPoint = p[(i + 1) % n];
        const V3 ttc = p[i] + nextPoint + centrePoint;
        const double ta = mag(cross(p[i] - centrePoint, nextPoint - centrePoint));
        sumA += ta;
        sumAc += ta * ttc;
    }
    if (sumA > VSMALL) return sumAc / (3.0 * sumA);
    return centrePoint;
}

// triPointRef::areaNormal
static inline V3 triAreaNormal(V3 a, V3 b, V3 c) { return 0.5 * cross(b - a, c - a); }

// face::areaNormal
static V3 faceAreaNormal(const std::vector<V3>& p) {
    const int n = (int)p.size();
    if (n == 3) return triAreaNormal(p[0], p[1], p[2]);
    V3 centrePoint = VZERO;
    for (int i = 0; i < n; ++i) centrePoint += p[i];
    centrePoint = centrePoint / (double)n;
    V3 nrm = VZERO;
    for (int i = 0; i < n; ++i) {
        const V3 nextPoint = (i < n - 1) ? p[i + 1] : p[0];
        nrm += triAreaNormal(p[i], nextPoint, centrePoint);
    }
    return nrm;
}

// primitiveMeshTools::makeFaceCentresAndAreas for one face
static void primitiveFaceGeom(const std::vector<V3>& p, V3& fc, V3& fa) {
    const int n = (int)p.size();
    if (n == 3) {
        fc = (1.0 / 3.0) * (p[0] + p[1] + p[2]);
        fa = 0.5 * cross(p[1] - p[0], p[2] - p[0]);
        return;
    }
    V3 sumN = VZERO;
    double sumA = 0.0;
    V3 sumAc = VZERO;
    V3 fCentre = p[0];
    for (int i = 1; i < n; ++i) fCentre += p[i];
    fCentre = fCentre / (double)n;
    for (int i = 0; i < n; ++i) {
        const V3 nextPoint = p[i == n - 1 ? 0 : i + 1];
        const V3 thisPoint = p[i];
        const V3 c = thisPoint + nextPoint + fCentre;
        const V3 nn = cross(nextPoint - thisPoint, fCentre - thisPoint);
        const double a = mag(nn);
        sumN += nn;
        sumA += a;
        sumAc += a * c;
    }
    if (sumA < ROOTVSMALL) {
        fc = fCentre;
        fa = VZERO;
    } else {
        fc = (1.0 / 3.0) * sumAc / sumA;
        fa = 0.5 * sumN;
    }
}

// ---------------------------------------------------------------------------
// Mesh: raw polyMesh + OpenFOAM demand-driven addressing and geometry
// ---------------------------------------------------------------------------
struct Mesh {
    int nP = 0, nF = 0, nIF = 0, nC = 0;
    std::vector<V3> points;
    std::vector<int> fOff, fPts, own, nei;
    std::vector<int> patchStart, patchSize, patchType;
    std::vector<int> bPatch;                   // patch of boundary face (f - nIF)
    // primitiveMesh::cells(): owned faces ascending, then neighbour faces ascending
    std::vector<std::vector<int>> cells;
    std::vector<std::vector<int>> cellPoints;  // ascending point labels
    std::vector<std::vector<int>> pointCells;  // ascending cell labels
    std::vector<V3> Cf, Sf, C;
    std::vector<double> magSf, V, flatness;
    // zoneCPCStencil-like stencil: the cell first, then every other cell sharing
    // a point with it, ascending
    std::vector<std::vector<int>> stencil;
    // volPointInterpolation weights
    std::vector<char> isPatchPoint;
    std::vector<std::vector<double>> pointW;          // internal points over pointCells
    std::vector<std::vector<int>> bPointFaces;        // boundary points: boundary faces (global)
    std::vector<std::vector<double>> bPointW;

    std::vector<V3> facePts(int f) const {
        std::vector<V3> r;
        r.reserve(fOff[f + 1] - fOff[f]);
        for (int k = fOff[f]; k < fOff[f + 1]; ++k) r.push_back(points[fPts[k]]);
        return r;
    }
    int faceSize(int f) const { return fOff[f + 1] - fOff[f]; }
    bool isInternal(int f) const { return f < nIF; }
    bool boundaryEmpty(int f) const {
        const int p = bPatch[f - nIF];
        return patchType[p] == 1 || patchSize[p] == 0;
    }
};

static void buildMesh(Mesh& m) {
    // cells() -- primitiveMesh::calcCells
    m.cells.assign(m.nC, {});
    for (int f = 0; f < m.nF; ++f) m.cells[m.own[f]].push_back(f);
    for (int f = 0; f < m.nIF; ++f) m.cells[m.nei[f]].push_back(f);
    // boundary patch ids
    m.bPatch.assign(m.nF - m.nIF, -1);
    for (size_t p = 0; p < m.patchStart.size(); ++p)
        for (int k = 0; k < m.patchSize[p]; ++k) m.bPatch[m.patchStart[p] + k - m.nIF] = (int)p;
    // pointCells (ascending cells) and cellPoints (ascending points)
    m.pointCells.assign(m.nP, {});
    for (int c = 0; c < m.nC; ++c) {
        std::vector<int> pts;
        for (int f : m.cells[c])
            for (int k = m.fOff[f]; k < m.fOff[f + 1]; ++k) pts.push_back(m.fPts[k]);
        std::sort(pts.begin(), pts.end());
        pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
        for (int p : pts) m.pointCells[p].push_back(c);
    }
    m.cellPoints.assign(m.nC, {});
    for (int p = 0; p < m.nP; ++p)
        for (int c : m.pointCells[p]) m.cellPoints[c].push_back(p);
    // face geometry
    m.Cf.resize(m.nF);
    m.Sf.resize(m.nF);
    m.magSf.resize(m.nF);
    for (int f = 0; f < m.nF; ++f) {
        primitiveFaceGeom(m.facePts(f), m.Cf[f], m.Sf[f]);
        m.magSf[f] = mag(m.Sf[f]);
    }
    // cell geometry -- primitiveMeshTools::makeCellCentresAndVols
    std::vector<V3> cEst(m.nC, VZERO), cellCtrs(m.nC, VZERO);
    std::vector<double> cellVols(m.nC, 0.0);
    std::vector<int> nCellFaces(m.nC, 0);
    for (int f = 0; f < m.nF; ++f) { cEst[m.own[f]] += m.Cf[f]; ++nCellFaces[m.own[f]]; }
    for (int f = 0; f < m.nIF; ++f) { cEst[m.nei[f]] += m.Cf[f]; ++nCellFaces[m.nei[f]]; }
    for (int c = 0; c < m.nC; ++c) cEst[c] = cEst[c] / (double)nCellFaces[c];
    for (int f = 0; f < m.nF; ++f) {
        const int o = m.own[f];
        const double pyr3Vol = dot(m.Sf[f], m.Cf[f] - cEst[o]);
        const V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[o];
        cellCtrs[o] += pyr3Vol * pc;
        cellVols[o] += pyr3Vol;
    }
    for (int f = 0; f < m.nIF; ++f) {
        const int n = m.nei[f];
        const double pyr3Vol = dot(m.Sf[f], cEst[n] - m.Cf[f]);
        const V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[n];
        cellCtrs[n] += pyr3Vol * pc;
        cellVols[n] += pyr3Vol;
    }
    m.C.resize(m.nC);
    m.V.resize(m.nC);
    for (int c = 0; c < m.nC; ++c) {
        if (std::fabs(cellVols[c]) > VSMALL) m.C[c] = cellCtrs[c] / cellVols[c];
        else m.C[c] = cEst[c];
        m.V[c] = cellVols[c] * (1.0 / 3.0);
    }
    // face flatness -- reconstruction.C:408-440
    m.flatness.assign(m.nF, 1.0);
    for (int f = 0; f < m.nF; ++f) {
        const int n = m.faceSize(f);
        if (n > 3 && m.magSf[f] > ROOTVSMALL) {
            const V3& fc = m.Cf[f];
            double sumA = 0.0;
            for (int i = 0; i < n; ++i) {
                const V3& thisPoint = m.points[m.fPts[m.fOff[f] + i]];
                const V3& nextPoint = m.points[m.fPts[m.fOff[f] + (i + 1) % n]];
                const V3 nn = 0.5 * cross(nextPoint - thisPoint, fc - thisPoint);
                sumA += mag(nn);
            }
            m.flatness[f] = m.magSf[f] / (sumA + ROOTVSMALL);
        }
    }
    // cell-point-cell stencil
    m.stencil.assign(m.nC, {});
    for (int c = 0; c < m.nC; ++c) {
        std::vector<int> s;
        for (int p : m.cellPoints[c])
            for (int cc : m.pointCells[p])
                if (cc != c) s.push_back(cc);
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        m.stencil[c].push_back(c);
        m.stencil[c].insert(m.stencil[c].end(), s.begin(), s.end());
    }
    // volPointInterpolation weights (makeInternalWeights / makeBoundaryWeights)
    m.isPatchPoint.assign(m.nP, 0);
    m.bPointFaces.assign(m.nP, {});
    for (int f = m.nIF; f < m.nF; ++f) {
        if (m.boundaryEmpty(f)) continue;
        for (int k = m.fOff[f]; k < m.fOff[f + 1]; ++k) {
            m.isPatchPoint[m.fPts[k]] = 1;
            m.bPointFaces[m.fPts[k]].push_back(f);
        }
    }
    m.pointW.assign(m.nP, {});
    m.bPointW.assign(m.nP, {});
    for (int p = 0; p < m.nP; ++p) {
        double sumW = 0.0;
        if (!m.isPatchPoint[p]) {
            for (int c : m.pointCells[p]) {
                const double w = 1.0 / mag(m.points[p] - m.C[c]);
                m.pointW[p].push_back(w);
                sumW += w;
            }
            for (double& w : m.pointW[p]) w /= sumW;
        } else {
            for (int f : m.bPointFaces[p]) {
                const double w = 1.0 / mag(m.points[p] - m.Cf[f]);
                m.bPointW[p].push_back(w);
                sumW += w;
            }
            for (double& w : m.bPointW[p]) w /= sumW;
        }
    }
}

// ---------------------------------------------------------------------------
// cutFace (src/SimPLIC/cut/cutFace/cutFace.C)
// ---------------------------------------------------------------------------
struct CutFace {
    const Mesh* m = nullptr;
    V3 subFaceCentre = VZERO, subFaceArea = VZERO;
    std::vector<V3> subFacePoints, interfacePoints;
    std::vector<double> pointDistances;
    int faceStatus = -1;

    void clearStorage() {
        subFaceCentre = VZERO;
        subFaceArea = VZERO;
        subFacePoints.clear();
        interfacePoints.clear();
        pointDistances.clear();
        faceStatus = -1;
    }

    // cutFace.C:37-96
    void calcSubFaceCentreAndArea() {
        const int nPoints = (int)subFacePoints.size();
        if (nPoints == 3) {
            subFaceCentre = (1.0 / 3.0) * (subFacePoints[0] + subFacePoints[1] + subFacePoints[2]);
            subFaceArea = 0.5 * cross(subFacePoints[1] - subFacePoints[0],
                                      subFacePoints[2] - subFacePoints[0]);
        } else {
            V3 sumN = VZERO;
            double sumA = 0.0;
            V3 sumAc = VZERO;
            V3 fCentre = subFacePoints[0];
            for (int pi = 1; pi < nPoints; pi++) fCentre += subFacePoints[pi];
            fCentre = fCentre / (double)nPoints;
            for (int pi = 0; pi < nPoints; pi++) {
                const V3& nextPoint = subFacePoints[(pi + 1) % nPoints];
                const V3 c = subFacePoints[pi] + nextPoint + fCentre;
                const V3 n = cross(nextPoint - subFacePoints[pi], fCentre - subFacePoints[pi]);
                const double a = mag(n);
                sumN += n;
                sumA += a;
                sumAc += a * c;
            }
            if (sumA < ROOTVSMALL) {
                subFaceCentre = fCentre;
                subFaceArea = VZERO;
            } else {
                subFaceCentre = (1.0 / 3.0) * sumAc / sumA;
                subFaceArea = 0.5 * sumN;
            }
        }
    }

    // cutFace.C:136-259
    int calcSubFace(const std::vector<V3>& fPts, V3 normal, double distance) {
        clearStorage();
        const int fsize = (int)fPts.size();
        int nSubmergedPoints = 0;
        int firstSubmergedPoint = -1;
        for (int i = 0; i < fsize; ++i) {
            double distanceI = dot(fPts[i], normal) + distance;
            if (std::fabs(distanceI) < TSMALL) distanceI += sgn(distanceI) * TSMALL;
            pointDistances.push_back(distanceI);
            if (distanceI < 0.0) {
                nSubmergedPoints++;
                if (firstSubmergedPoint == -1) firstSubmergedPoint = i;
            }
        }
        if (nSubmergedPoints == fsize) {
            faceStatus = -1;
            subFaceCentre = faceCentre(fPts);
            subFaceArea = faceAreaNormal(fPts);
            return faceStatus;
        } else if (nSubmergedPoints == 0) {
            faceStatus = 1;
            subFaceCentre = VZERO;
            subFaceArea = VZERO;
            return faceStatus;
        }
        faceStatus = 0;
        for (int i = firstSubmergedPoint; i < firstSubmergedPoint + fsize; ++i) {
            const int currentId = i % fsize;
            const int nextId = (i + 1) % fsize;
            if (pointDistances[currentId] < 0) subFacePoints.push_back(fPts[currentId]);
            if ((pointDistances[currentId] * pointDistances[nextId]) < 0) {
                const double weight = pointDistances[currentId] /
                                      (pointDistances[currentId] - pointDistances[nextId]);
                const V3 cutPoint = fPts[currentId] + weight * (fPts[nextId] - fPts[currentId]);
                subFacePoints.push_back(cutPoint);
                interfacePoints.push_back(cutPoint);
            }
        }
        if (subFacePoints.size() >= 3) {
            faceStatus = 0;
            calcSubFaceCentreAndArea();
        } else {
            faceStatus = -1;
            subFaceCentre = faceCentre(fPts);
            subFaceArea = faceAreaNormal(fPts);
        }
        return faceStatus;
    }

    int calcSubFace(int faceI, V3 normal, double distance) {
        return calcSubFace(m->facePts(faceI), normal, distance);
    }

    // cutFace.C:392-506
    double timeIntegratedArea(const std::vector<V3>& fPts, V3 normal, double distance,
                              const std::vector<double>& pTimes, double Un0, double dt,
                              double magSf) {
        double tIntArea = 0.0;
        std::vector<int> order(pTimes.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return pTimes[a] < pTimes[b]; });
        const double firstTime = pTimes[order.front()];
        const double lastTime = pTimes[order.back()];
        if (lastTime <= 0.0) {
            tIntArea = magSf * dt * pos0(Un0);
            return tIntArea;
        }
        if (firstTime >= dt) {
            tIntArea = magSf * dt * (1.0 - pos0(Un0));
            return tIntArea;
        }
        std::vector<double> sortedTimes;
        double prevTime = 0.0;
        double subAreaOld = 0.0, subAreaNew = 0.0, subAreaMid = 0.0;
        if (firstTime > 0.0) {
            subAreaOld = magSf * (1.0 - pos0(Un0));
            tIntArea = subAreaOld * firstTime;
            sortedTimes.push_back(firstTime);
            prevTime = firstTime;
        } else {
            sortedTimes.push_back(0.0);
            prevTime = 0.0;
            calcSubFace(fPts, normal, distance);
            subAreaOld = mag(subFaceArea);
        }
        const double smallTime = std::max(TSMALL / std::fabs(Un0), TSMALL);
        for (size_t ti = 0; ti < order.size(); ++ti) {
            const double timeI = pTimes[order[ti]];
            if (timeI > (prevTime + smallTime) && timeI < dt) {
                sortedTimes.push_back(timeI);
                prevTime = timeI;
            }
        }
        if (lastTime > dt) {
            sortedTimes.push_back(dt);
        } else {
            tIntArea += magSf * (dt - lastTime) * pos0(Un0);
        }
        for (int k = 0; k < (int)sortedTimes.size() - 1; k++) {
            const double tauOld = sortedTimes[k];
            const double tauNew = sortedTimes[k + 1];
            const double deltaTau = 0.5 * (tauNew - tauOld);
            calcSubFace(fPts, normal, distance - tauNew * Un0);
            subAreaNew = mag(subFaceArea);
            calcSubFace(fPts, normal, distance - (tauOld + deltaTau) * Un0);
            subAreaMid = mag(subFaceArea);
            tIntArea += (deltaTau / 3.0) * (subAreaOld + 4.0 * subAreaMid + subAreaNew);
            subAreaOld = subAreaNew;
        }
        return tIntArea;
    }

    // cutFace.C:262-389
    double timeIntegratedFaceFlux(int faceI, V3 normal, double distance, double Un0, double dt,
                                  double phi, double magSf) {
        if (std::fabs(phi) <= TSMALL) return 0.0;
        const std::vector<V3> fPts = m->facePts(faceI);
        const int nPoints = (int)fPts.size();
        if (std::fabs(Un0 * dt) > TSMALL) {
            std::vector<double> pTimes(nPoints);
            for (int i = 0; i < nPoints; ++i) {
                const double pTimeI = (dot(fPts[i], normal) + distance) / Un0;
                pTimes[i] = std::fabs(pTimeI) < TSMALL ? 0.0 : pTimeI;
            }
            double dVf = 0.0;
            if (m->flatness[faceI] > (1.0 - TSMALL)) {
                dVf = phi / magSf *
                      timeIntegratedArea(fPts, normal, distance, pTimes, Un0, dt, magSf);
            } else {
                std::vector<V3> fPtsTri(3);
                std::vector<double> pTimesTri(3);
                fPtsTri[0] = m->Cf[faceI];
                const double pTimeTri0 = (dot(fPtsTri[0], normal) + distance) / Un0;
                pTimesTri[0] = std::fabs(pTimeTri0) < TSMALL ? 0.0 : pTimeTri0;
                for (int pi = 0; pi < nPoints; ++pi) {
                    fPtsTri[1] = fPts[pi];
                    pTimesTri[1] = pTimes[pi];
                    fPtsTri[2] = fPts[(pi + 1) % nPoints];
                    pTimesTri[2] = pTimes[(pi + 1) % nPoints];
                    const double magSfTri = mag(triAreaNormal(fPtsTri[0], fPtsTri[1], fPtsTri[2]));
                    const double phiTri = phi * magSfTri / magSf;
                    dVf += phiTri / magSfTri *
                           timeIntegratedArea(fPtsTri, normal, distance, pTimesTri, Un0, dt,
                                              magSfTri);
                }
            }
            return dVf;
        } else {
            if (m->flatness[faceI] > (1.0 - TSMALL)) {
                calcSubFace(fPts, normal, distance);
                const double alphaf = mag(subFaceArea) / magSf;
                return (phi * dt * alphaf);
            } else {
                std::vector<V3> fPtsTri(3);
                fPtsTri[0] = m->Cf[faceI];
                double dVf = 0.0;
                for (int pi = 0; pi < nPoints; ++pi) {
                    fPtsTri[1] = fPts[pi];
                    fPtsTri[2] = fPts[(pi + 1) % nPoints];
                    const double magSfTri = mag(triAreaNormal(fPtsTri[0], fPtsTri[1], fPtsTri[2]));
                    const double phiTri = phi * magSfTri / magSf;
                    calcSubFace(fPtsTri, normal, distance);
                    const double alphafTri = mag(subFaceArea) / magSfTri;
                    dVf += (phiTri * dt * alphafTri);
                }
                return dVf;
            }
        }
    }
};

// ---------------------------------------------------------------------------
// cutCell (src/SimPLIC/cut/cutCell/cutCell.C)
// ---------------------------------------------------------------------------
struct CutCell {
    const Mesh* m = nullptr;
    CutFace cutFace;
    int cellStatus = -1;
    std::vector<V3> cutFaceCentres, cutFaceAreas;
    std::vector<std::vector<V3>> interfaceEdges;
    V3 subCellCentre = VZERO;
    double subCellVolume = -10.0, VOF = -10.0;
    V3 interfaceCentre = VZERO, interfaceArea = VZERO;
    int cellI = -1;
    // face-cut classification of the last calcSubCell: 2 bits per face
    // (status + 1), and the number of plane crossings on its cut faces
    uint64_t faceClass = 0;
    int nCutPoints = 0;
    // split-warped-face local description (cutCell.C:140-236)
    std::vector<V3> localPoints;
    std::vector<std::vector<int>> localFaces;

    void clearStorage() {
        cellI = -1;
        cellStatus = -1;
        cutFaceCentres.clear();
        cutFaceAreas.clear();
        interfaceEdges.clear();
        interfaceCentre = VZERO;
        interfaceArea = VZERO;
        subCellCentre = VZERO;
        subCellVolume = -10.0;
        VOF = -10.0;
    }

    // cutCell.C:37-100
    void calcInterfaceCentreAndArea() {
        V3 fCentre = VZERO;
        int nEdgePoints = 0;
        for (const auto& edgePoints : interfaceEdges)
            for (const V3& p : edgePoints) { fCentre += p; nEdgePoints++; }
        if (nEdgePoints > 0) fCentre = fCentre / (double)nEdgePoints;
        V3 sumN = VZERO;
        double sumA = 0.0;
        V3 sumAc = VZERO;
        for (const auto& edgePoints : interfaceEdges) {
            const int nPoints = (int)edgePoints.size();
            for (int pi = 0; pi < nPoints - 1; pi++) {
                const V3& nextPoint = edgePoints[pi + 1];
                const V3 c = edgePoints[pi] + nextPoint + fCentre;
                const V3 n = cross(nextPoint - edgePoints[pi], fCentre - edgePoints[pi]);
                const double a = mag(n);
                sumN += sgn(dot(n, sumN)) * n;
                sumA += a;
                sumAc += a * c;
            }
        }
        if (sumA < ROOTVSMALL) {
            interfaceCentre = fCentre;
            interfaceArea = VZERO;
        } else {
            interfaceCentre = (1.0 / 3.0) * sumAc / sumA;
            interfaceArea = 0.5 * sumN;
        }
        // orientation test against subCellCentre, which is still zero here
        if (dot(interfaceArea, interfaceCentre - subCellCentre) < 0.0)
            interfaceArea = (-1.0) * interfaceArea;
    }

    // cutCell.C:103-137
    void calcSubCellCentreAndVolume() {
        subCellCentre = VZERO;
        subCellVolume = 0.0;
        V3 sum = VZERO;
        for (const V3& c : cutFaceCentres) sum += c;
        const V3 cEst = sum / (double)cutFaceCentres.size();
        for (size_t f = 0; f < cutFaceCentres.size(); ++f) {
            const double pyr3Vol =
                std::max(std::fabs(dot(cutFaceAreas[f], cutFaceCentres[f] - cEst)), VSMALL);
            const V3 pc = 0.75 * cutFaceCentres[f] + 0.25 * cEst;
            subCellCentre += pyr3Vol * pc;
            subCellVolume += pyr3Vol;
        }
        subCellCentre = subCellCentre / subCellVolume;
        subCellVolume /= 3.0;
    }

    // cutCell.C:140-236 (only the split branch is needed: the unsplit branch
    // reproduces the mesh faces, which calcSubCell reads directly)
    void getLocalPointFieldAndFaceList(int cell) {
        localPoints.clear();
        localFaces.clear();
        // c.labels(faces): unique point labels in order of first appearance over the cell faces
        std::vector<int> globalPointLabels;
        for (int f : m->cells[cell])
            for (int k = m->fOff[f]; k < m->fOff[f + 1]; ++k) {
                const int p = m->fPts[k];
                if (std::find(globalPointLabels.begin(), globalPointLabels.end(), p) ==
                    globalPointLabels.end())
                    globalPointLabels.push_back(p);
            }
        for (int p : globalPointLabels) localPoints.push_back(m->points[p]);
        auto findLocal = [&](int p) {
            return (int)(std::find(globalPointLabels.begin(), globalPointLabels.end(), p) -
                         globalPointLabels.begin());
        };
        auto reverseFace = [](const std::vector<int>& f) {
            // OpenFOAM face::reverseFace keeps the first point
            std::vector<int> r(f.size());
            r[0] = f[0];
            for (size_t i = 1; i < f.size(); ++i) r[i] = f[f.size() - i];
            return r;
        };
        for (int f : m->cells[cell]) {
            const int n = m->faceSize(f);
            if (m->flatness[f] > (1.0 - TSMALL)) {
                std::vector<int> localFa;
                for (int k = 0; k < n; ++k) localFa.push_back(findLocal(m->fPts[m->fOff[f] + k]));
                localFaces.push_back(cell == m->own[f] ? localFa : reverseFace(localFa));
            } else {
                localPoints.push_back(m->Cf[f]);
                for (int k = 0; k < n; ++k) {
                    const int nk = (k + 1) % n;
                    std::vector<int> localFa(3);
                    localFa[0] = (int)localPoints.size() - 1;
                    localFa[1] = findLocal(m->fPts[m->fOff[f] + k]);
                    localFa[2] = findLocal(m->fPts[m->fOff[f] + nk]);
                    localFaces.push_back(cell == m->own[f] ? localFa : reverseFace(localFa));
                }
            }
        }
    }

    void accumulateFace(int faceStatus, const std::vector<V3>& fPts, bool& fullySubmerged,
                        bool& fullyEmpty, int& nSubmergedFaces) {
        (void)fPts;
        if (faceStatus == 0) {
            cutFaceCentres.push_back(cutFace.subFaceCentre);
            cutFaceAreas.push_back(cutFace.subFaceArea);
            interfaceEdges.push_back(cutFace.interfacePoints);
            fullySubmerged = false;
            fullyEmpty = false;
        } else if (faceStatus == -1) {
            cutFaceCentres.push_back(cutFace.subFaceCentre);
            cutFaceAreas.push_back(cutFace.subFaceArea);
            fullyEmpty = false;
            nSubmergedFaces++;
        } else {
            fullySubmerged = false;
        }
    }

    // Vector::normalise of OpenFOAM: zero below ROOTVSMALL, else divided by
    // its magnitude component by component
    static V3 normalised(V3 v) {
        const double s = mag(v);
        return s < ROOTVSMALL ? VZERO : v / s;
    }

    // cutCell.C:545-608: the plane crossings of the cut faces of the last
    // calcSubCell, ordered by angle about the interface centre in a frame
    // with z along the interface area vector and x towards the first
    // crossing; a crossing whose angle lies within 1e-8 of the previous one
    // in that order is dropped.  sortedOrder is a stable sort.
    std::vector<V3> interfacePoints() const {
        std::vector<V3> out;
        if (cellStatus != 0) return out;
        const V3 zhat = interfaceArea / mag(interfaceArea);
        V3 xhat = interfaceEdges[0][0] - interfaceCentre;
        xhat = xhat - dot(xhat, zhat) * zhat;
        xhat = normalised(xhat);
        const V3 yhat = normalised(cross(zhat, xhat));
        std::vector<V3> pts;
        std::vector<double> ang;
        for (const auto& edgePoints : interfaceEdges)
            for (const V3& p : edgePoints) {
                pts.push_back(p);
                ang.push_back(std::atan2(dot(p - interfaceCentre, yhat), dot(p - interfaceCentre, xhat)));
            }
        std::vector<int> order(pts.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ang[a] < ang[b]; });
        out.push_back(pts[order[0]]);
        for (size_t k = 1; k < order.size(); ++k)
            if (std::fabs(ang[order[k]] - ang[order[k - 1]]) > 1e-8) out.push_back(pts[order[k]]);
        return out;
    }

    // cutCell.C:343-542
    int calcSubCell(int cell, V3 normal, double distance, bool splitWarpedFace) {
        clearStorage();
        cellI = cell;
        bool fullySubmerged = true, fullyEmpty = true;
        int nSubmergedFaces = 0;
        faceClass = 0;
        nCutPoints = 0;
        int faceNo = 0;
        auto classify = [&](int st) {
            faceClass |= (uint64_t)(st + 1) << (2 * faceNo);
            if (st == 0) nCutPoints += (int)cutFace.interfacePoints.size();
            faceNo++;
        };
        if (splitWarpedFace) {
            for (const auto& lf : localFaces) {
                std::vector<V3> fPts;
                for (int q : lf) fPts.push_back(localPoints[q]);
                const int st = cutFace.calcSubFace(fPts, normal, distance);
                classify(st);
                accumulateFace(st, fPts, fullySubmerged, fullyEmpty, nSubmergedFaces);
            }
        } else {
            for (int f : m->cells[cell]) {
                const std::vector<V3> fPts = m->facePts(f);
                const int st = cutFace.calcSubFace(fPts, normal, distance);
                classify(st);
                accumulateFace(st, fPts, fullySubmerged, fullyEmpty, nSubmergedFaces);
            }
        }
        if (!fullySubmerged && !fullyEmpty) {
            cellStatus = 0;
            calcInterfaceCentreAndArea();
            if (mag(interfaceArea) < TSMALL) {
                if (nSubmergedFaces == 0) {
                    cellStatus = 1;
                    subCellCentre = VZERO;
                    subCellVolume = 0.0;
                    VOF = 0.0;
                    return cellStatus;
                } else {
                    cellStatus = -1;
                    subCellCentre = m->C[cell];
                    subCellVolume = m->V[cell];
                    VOF = 1.0;
                    return cellStatus;
                }
            }
            cutFaceCentres.push_back(interfaceCentre);
            cutFaceAreas.push_back(interfaceArea);
            calcSubCellCentreAndVolume();
            VOF = subCellVolume / m->V[cell];
        } else if (fullyEmpty) {
            cellStatus = 1;
            subCellCentre = VZERO;
            subCellVolume = 0.0;
            VOF = 0.0;
        } else if (fullySubmerged) {
            cellStatus = -1;
            subCellCentre = m->C[cell];
            subCellVolume = m->V[cell];
            VOF = 1.0;
        }
        return cellStatus;
    }

    // cutCell.C:611-799.  Writes d, C, S only on the paths where the
    // reference writes interfaceD/C/S; returns the cell status.
    int findSignedDistance(int cell, double alphaI, V3 normalI, bool splitWarpedFace, double& dOut,
                           V3& cOut, V3& sOut, int* nCutsOut = nullptr) {
        cellI = cell;
        int nCuts = 0;
        faceClass = 0;
        nCutPoints = 0;
        if (mag(normalI) < TSMALL) {
            if (nCutsOut) *nCutsOut = 0;
            return (int)sgn(0.5 - alphaI);
        }
        std::vector<double> vertexDistances;
        if (splitWarpedFace) {
            getLocalPointFieldAndFaceList(cell);
            for (const V3& p : localPoints) vertexDistances.push_back(-dot(normalI, p));
        } else {
            for (int p : m->cellPoints[cell]) vertexDistances.push_back(-dot(normalI, m->points[p]));
        }
        std::vector<int> order(vertexDistances.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return vertexDistances[a] > vertexDistances[b]; });
        double lowDistance = vertexDistances[order.front()];
        double upDistance = vertexDistances[order.back()];
        int lowLabel = 0;
        int upLabel = (int)vertexDistances.size() - 1;
        double lowAlpha = 0.0, upAlpha = 1.0;
        while ((upLabel - lowLabel) > 1) {
            const double midLabelS = std::round(0.5 * (upLabel + lowLabel));
            const int midLabel = (int)midLabelS;
            const double midDistance = vertexDistances[order[midLabel]];
            calcSubCell(cell, normalI, midDistance, splitWarpedFace);
            nCuts++;
            const double midAlpha = VOF;
            if (std::fabs(midAlpha - alphaI) < TSMALL) {
                dOut = midDistance;
                cOut = interfaceCentre;
                sOut = interfaceArea;
                if (nCutsOut) *nCutsOut = nCuts;
                return cellStatus;
            }
            if (midAlpha > alphaI) {
                upLabel = midLabel;
                upDistance = midDistance;
                upAlpha = midAlpha;
            } else {
                lowLabel = midLabel;
                lowDistance = midDistance;
                lowAlpha = midAlpha;
            }
        }
        if (std::fabs(lowDistance - upDistance) < TSMALL) {
            const double midDistance = 0.5 * (lowDistance + upDistance);
            calcSubCell(cell, normalI, midDistance, splitWarpedFace);
            nCuts++;
            dOut = midDistance;
            cOut = interfaceCentre;
            sOut = interfaceArea;
            if (nCutsOut) *nCutsOut = nCuts;
            return cellStatus;
        }
        const double alphaPrismatoid = upAlpha - lowAlpha;
        const double deltaDistance = (upDistance - lowDistance) / 3.0;
        const double distanceOneThird = lowDistance + deltaDistance;
        calcSubCell(cell, normalI, distanceOneThird, splitWarpedFace);
        nCuts++;
        const double alphaOneThird = VOF - lowAlpha;
        const double distanceTwoThirds = lowDistance + 2.0 * deltaDistance;
        calcSubCell(cell, normalI, distanceTwoThirds, splitWarpedFace);
        nCuts++;
        const double alphaTwoThirds = VOF - lowAlpha;
        const double a = 13.5 * alphaOneThird - 13.5 * alphaTwoThirds + 4.5 * alphaPrismatoid;
        const double b = -22.5 * alphaOneThird + 18.0 * alphaTwoThirds - 4.5 * alphaPrismatoid;
        const double c = 9.0 * alphaOneThird - 4.5 * alphaTwoThirds + 1.0 * alphaPrismatoid;
        const double d = lowAlpha - alphaI;
        double lambda = 0.5;
        for (int iter = 0; iter < 100; iter++) {
            const double func = a * (lambda * lambda * lambda) + b * (lambda * lambda) + c * lambda + d;
            const double funcPrime = 3.0 * a * (lambda * lambda) + 2.0 * b * lambda + c;
            const double lambdaNew = lambda - (func / funcPrime);
            if (std::fabs(lambdaNew - lambda) < TSMALL) break;
            lambda = lambdaNew;
        }
        const double distance0 = lowDistance - lambda * (lowDistance - upDistance);
        calcSubCell(cell, normalI, distance0, splitWarpedFace);
        nCuts++;
        dOut = distance0;
        cOut = interfaceCentre;
        sOut = interfaceArea;
        if (nCutsOut) *nCutsOut = nCuts;
        return cellStatus;
    }
};

// ---------------------------------------------------------------------------
// OpenFOAM scalarMatrices: Cholesky "LUDecompose" of a symmetric matrix and
// invDecomposed; used by the polyDegree1 least-squares gradient.
// ---------------------------------------------------------------------------
static void choleskyDecompose(double A[4][4]) {
    const int size = 4;
    for (int j = 0; j < size; j++)
        for (int k = j + 1; k < size; k++) A[j][k] = 0.0;
    for (int j = 0; j < size; j++) {
        double d = 0.0;
        for (int k = 0; k < j; k++) {
            double s = 0.0;
            for (int i = 0; i < k; i++) s += A[i][k] * A[i][j];
            s = (A[j][k] - s) / A[k][k];
            A[k][j] = s;
            A[j][k] = s;
            d += s * s;
        }
        d = A[j][j] - d;
        A[j][j] = std::sqrt(d);
    }
}

static void invDecomposed(const double L[4][4], double R[4][4]) {
    const int n = 4;
    double inv[4][4] = {};
    for (int i = 0; i < n; ++i) {
        inv[i][i] = 1.0 / L[i][i];
        for (int j = 0; j < i; ++j) {
            double sum = 0.0;
            for (int k = j; k < i; k++) sum -= L[i][k] * inv[k][j];
            inv[i][j] = sum / L[i][i];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) R[i][j] = 0.0;
    for (int k = 0; k < n; k++)
        for (int i = 0; i <= k; i++)
            for (int j = 0; j <= k; j++) R[i][j] += inv[k][i] * inv[k][j];
}

// leastSquareGrad<scalar>("polyDegree1").grad(positions, values): terms
// (1, dx, dy, dz), normal equations accumulated in stencil order.
static V3 lsGrad(const std::vector<V3>& pos, const std::vector<double>& val) {
    double A[4][4] = {};
    double B[4] = {};
    for (size_t s = 0; s < pos.size(); ++s) {
        const double t[4] = {1.0, pos[s].x, pos[s].y, pos[s].z};
        const double w = g_mutation == 1 ? 1.0 / (magSqr(pos[s]) + 1e-3) : 1.0;
        for (int r = 0; r < 4; ++r) {
            for (int c = 0; c < 4; ++c) A[r][c] += g_mutation == 1 ? w * t[r] * t[c] : t[r] * t[c];
            B[r] += g_mutation == 1 ? w * t[r] * val[s] : t[r] * val[s];
        }
    }
    double invA[4][4];
    choleskyDecompose(A);
    invDecomposed(A, invA);
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) x[i] += invA[i][j] * B[j];
    return v3(x[1], x[2], x[3]);
}

// ---------------------------------------------------------------------------
// Solver state: reconstruction + advection (solveVofEqu)
// ---------------------------------------------------------------------------
struct Solver {
    const Mesh* m = nullptr;
    double mixedCellTol = 1e-8;
    int nAlphaBounds = 10;
    double snapTol = 0.0;
    bool clip = true;
    bool splitWarpedFace = false;
    int orientation = 0;  // 0: isoAlphaGrad (LS), 1: alphaGrad, 2: isoRDF
    double rdfTol = 1e-6, rdfRelTol = 0.1;
    int rdfIterations = 5;
    int nThreads = 1;
    // fields
    std::vector<double> alpha, alphaOld, phi, dVf, alphaPhi;
    std::vector<V3> U, Ub;
    // reconstruction outputs
    std::vector<int> mixedCells, cellStatus, nCuts, nCutPts;
    std::vector<uint64_t> faceClass;
    std::vector<V3> interfaceN, interfaceC, interfaceS;
    std::vector<double> interfaceD;
    std::vector<double> Un0;             // per mixed cell (0 if status != 0)
    double reconstructionTime = 0.0, advectionTime = 0.0;

    void init() {
        alpha.assign(m->nC, 0.0);
        alphaOld.assign(m->nC, 0.0);
        phi.assign(m->nF, 0.0);
        dVf.assign(m->nF, 0.0);
        alphaPhi.assign(m->nF, 0.0);
        U.assign(m->nC, VZERO);
        Ub.assign(m->nF - m->nIF, VZERO);
        interfaceN.assign(m->nC, VZERO);
        interfaceC.assign(m->nC, VZERO);
        interfaceS.assign(m->nC, VZERO);
        interfaceD.assign(m->nC, 0.0);
    }

    bool isAMixedCell(int c) const {
        return mixedCellTol < alpha[c] && alpha[c] < 1.0 - mixedCellTol;
    }

    // reconstruction.C:634-677
    void initialize() {
        mixedCells.clear();
        cellStatus.clear();
        std::fill(interfaceN.begin(), interfaceN.end(), VZERO);
        std::fill(interfaceD.begin(), interfaceD.end(), 0.0);
        std::fill(interfaceC.begin(), interfaceC.end(), VZERO);
        std::fill(interfaceS.begin(), interfaceS.end(), VZERO);
        for (int c = 0; c < m->nC; ++c)
            if (isAMixedCell(c)) {
                mixedCells.push_back(c);
                cellStatus.push_back(-100);
            }
    }

    // reconstruction.C:85-141
    void calcInterfaceNFromIsoAlphaGrad() {
        const int nMixed = (int)mixedCells.size();
#pragma omp parallel for num_threads(nThreads) schedule(dynamic, 64)
        for (int i = 0; i < nMixed; ++i) {
            const int celli = mixedCells[i];
            std::vector<V3> cellCentre;
            std::vector<double> alphaValues;
            for (int g : m->stencil[celli]) {
                cellCentre.push_back(m->C[g]);
                alphaValues.push_back(alpha[g]);
            }
            for (V3& p : cellCentre) p -= m->C[celli];
            interfaceN[celli] = -lsGrad(cellCentre, alphaValues);
        }
        // interfaceNIn /= (mag(interfaceNIn) + SMALL) over all cells; non-mixed
        // cells hold zero vectors, which stay zero.
        for (int i = 0; i < nMixed; ++i) {
            const int c = mixedCells[i];
            interfaceN[c] = interfaceN[c] / (mag(interfaceN[c]) + SMALL);
        }
    }

    // reconstruction.C:74-82: interfaceN = -fvc::grad(alpha1) with the
    // Gauss linear scheme, normalised.  fvc::grad restates OpenFOAM v2312's
    // gaussGrad::gradf (faces in ascending order: += Sf*ssf on the owner, -=
    // on the neighbour, then the boundary faces of non-empty patches, then
    // /= V) over linear interpolation with surfaceInterpolation::makeWeights
    // weights; OpenFOAM is a dependency of the reference, not part of it.
    // alpha is zeroGradient on the boundary.
    //
    // gradScheme 1 is Gauss pointLinear (fvSchemes grad(alpha1), selected by
    // tutorials/test/plicVofOrientationFoam/NAG/system/fvSchemes:35): the
    // linear face value plus OpenFOAM v2312's pointLinear correction.  The
    // face is fanned into triangles (pi, p_k, p_k-1), closing with (pi, p_0,
    // p_last), around pi = w C_own + (1 - w) C_nei, where pointLinear takes w
    // from the face weights at the index of the owner cell; each triangle adds
    // its area times the mean of the linear face value and its two
    // volPointInterpolation point values; the correction is that area average
    // minus the linear value, and the face value is linear + correction.
    // Boundary faces take no correction (non-coupled patches).
    int gradScheme = 0;

    double faceWeight(int f) const {
        const int o = m->own[f], n = m->nei[f];
        const double sfdOwn = std::fabs(dot(m->Sf[f], m->Cf[f] - m->C[o]));
        const double sfdNei = std::fabs(dot(m->Sf[f], m->C[n] - m->Cf[f]));
        return sfdNei / (sfdOwn + sfdNei);
    }

    // volPointInterpolation of alpha: internal points over their cells,
    // boundary points over their boundary faces, whose zeroGradient value is
    // the owner's alpha
    double pointAlpha(int p) const {
        double t = 0.0;
        if (!m->isPatchPoint[p]) {
            const auto& pc = m->pointCells[p];
            for (size_t i = 0; i < pc.size(); ++i) t += m->pointW[p][i] * alpha[pc[i]];
        } else {
            const auto& pf = m->bPointFaces[p];
            for (size_t i = 0; i < pf.size(); ++i) t += m->bPointW[p][i] * alpha[m->own[pf[i]]];
        }
        return t;
    }

    double pointLinearFaceValue(int f, double lin) const {
        const int o = m->own[f], n = m->nei[f];
        const double wq = g_mutation == 4 ? faceWeight(f) : faceWeight(o);
        const V3 pi = wq * m->C[o] + (1.0 - wq) * m->C[n];
        const int b = m->fOff[f], np = m->fOff[f + 1] - b;
        auto triMag = [&](V3 a, V3 bb, V3 c) { return mag(0.5 * cross(bb - a, c - a)); };
        const int p0 = m->fPts[b], pl = m->fPts[b + np - 1];
        // planted mistake 5: the closing triangle (pi, p_0, p_last) left out
        double at = g_mutation == 5 ? 0.0 : triMag(pi, m->points[p0], m->points[pl]);
        double sumAt = at;
        double sumPsip = at * (1.0 / 3.0) * (lin + pointAlpha(p0) + pointAlpha(pl));
        for (int k = 1; k < np; ++k) {
            const int pk = m->fPts[b + k], pk1 = m->fPts[b + k - 1];
            at = triMag(pi, m->points[pk], m->points[pk1]);
            sumAt += at;
            sumPsip += at * (1.0 / 3.0) * (lin + pointAlpha(pk) + pointAlpha(pk1));
        }
        const double corr = sumPsip / sumAt - lin;
        return lin + corr;
    }

    void calcInterfaceNFromRegAlphaGrad() {
        std::vector<V3> g(m->nC, VZERO);
        for (int f = 0; f < m->nIF; ++f) {
            const int o = m->own[f], n = m->nei[f];
            const double w = faceWeight(f);
            double ssf = w * (alpha[o] - alpha[n]) + alpha[n];
            if (gradScheme == 1) ssf = pointLinearFaceValue(f, ssf);
            const V3 v = ssf * m->Sf[f];
            g[o] += v;
            g[n] -= v;
        }
        for (int f = m->nIF; f < m->nF; ++f) {
            if (m->boundaryEmpty(f)) continue;
            g[m->own[f]] += alpha[m->own[f]] * m->Sf[f];
        }
        for (int c = 0; c < m->nC; ++c) {
            const V3 n = -(g[c] / m->V[c]);
            interfaceN[c] = n / (mag(n) + SMALL);
        }
    }

    // isoInterfaceGrad (reconstruction.C:144-191): the polyDegree1 least
    // squares gradient of a cell field over the cell-point-cell stencil of
    // every mixed cell (not normalised)
    void isoInterfaceGrad(const std::vector<double>& field, std::vector<V3>& g) const {
        const int nMixed = (int)mixedCells.size();
        for (int i = 0; i < nMixed; ++i) {
            const int celli = mixedCells[i];
            std::vector<V3> cellCentre;
            std::vector<double> values;
            for (int k : m->stencil[celli]) {
                cellCentre.push_back(m->C[k] - m->C[celli]);
                values.push_back(field[k]);
            }
            g[i] = lsGrad(cellCentre, values);
        }
    }

    // Vector::normalise(tol) of OpenFOAM, in place: zero below tol, else
    // divided by its magnitude component by component
    static V3& normaliseInPlace(V3& v, double tol) {
        const double s = mag(v);
        if (s < tol) v = VZERO;
        else v = v / s;
        return v;
    }

    // reconstruction.C:194-405 (isoRDF, also selected as RDF) with OpenFOAM
    // v2312's reconstructedDistanceFunction: markCellsNearSurf(mixed, 1) and
    // constructRDF.  OpenFOAM is a dependency of the reference, not part of
    // it; updateContactAngle changes nothing without contact-angle patches,
    // which this mesh description does not have.
    int rdfIterationsRun = 0;
    void calcInterfaceNFromIsoRDF() {
        const double TSMALL_ = 10.0 * SMALL;
        const int nMixed = (int)mixedCells.size();
        const int nC = m->nC;
        std::vector<char> isMixed(nC, 0), next(nC, 0);
        for (int c : mixedCells) isMixed[c] = 1;
        // level 0: interface cells, level 1: the cells sharing a point with one
        for (int c : mixedCells)
            for (int k : m->stencil[c]) next[k] = 1;
        std::vector<V3> g(nMixed, VZERO);
        isoInterfaceGrad(alpha, g);
        std::fill(interfaceC.begin(), interfaceC.end(), VZERO);
        std::fill(interfaceN.begin(), interfaceN.end(), VZERO);
        std::vector<char> tooCoarse(nC, 0);
        std::vector<double> rdf(nC, 0.0);
        std::vector<double> normalRes(nMixed), avgAngle(nMixed);
        rdfIterationsRun = 0;
        for (int iter = 0; iter < rdfIterations; ++iter) {
            rdfIterationsRun = iter + 1;
            for (int i = 0; i < nMixed; ++i) {
                const int c = mixedCells[i];
                interfaceN[c] = -normaliseInPlace(g[i], SMALL);
                if (tooCoarse[c]) continue;
                double d = interfaceD[c];
                V3 C = interfaceC[c], S = interfaceS[c];
                CutCell cc;
                cc.m = m;
                cc.cutFace.m = m;
                int nc = 0;
                cellStatus[i] = cc.findSignedDistance(c, alpha[c], interfaceN[c], splitWarpedFace, d, C, S, &nc);
                interfaceD[c] = d;
                interfaceC[c] = C;
                interfaceS[c] = S;
            }
            // constructRDF(nextToInterface, interfaceC, -interfaceN)
            for (int c = 0; c < nC; ++c) {
                if (!next[c]) {
                    rdf[c] = 0.0;
                    continue;
                }
                const V3 nrm = -interfaceN[c];
                if (mag(nrm) != 0.0) {
                    const V3 n = (-nrm) / mag(nrm);
                    rdf[c] = dot(interfaceC[c] - m->C[c], n);
                } else {
                    double averageWeight = 0.0, avgDist = 0.0;
                    for (int k : m->stencil[c]) {
                        V3 n = -(-interfaceN[k]);
                        if (mag(n) != 0.0) {
                            n = n / mag(n);
                            V3 dv = interfaceC[k] - m->C[c];
                            const double distToSurf = dot(dv, n);
                            double weight = 0.0;
                            if (mag(dv) != 0.0) {
                                dv = dv / mag(dv);
                                const double w = std::fabs(dot(dv, n));
                                weight = w * w;
                            } else {
                                weight = 1.0;
                            }
                            averageWeight += weight;
                            avgDist += distToSurf * weight;
                        }
                    }
                    if (averageWeight != 0.0) rdf[c] = avgDist / averageWeight;
                }
            }
            isoInterfaceGrad(rdf, g);
            for (int i = 0; i < nMixed; ++i) {
                const int c = mixedCells[i];
                // the test normalises g[i] in place only when the first
                // operand of the 'or' is false; its result is not used
                if (!(mag(interfaceN[c]) < TSMALL_)) (void)normaliseInPlace(g[i], SMALL);
                double avgDiffNormal = 0.0, weight = 0.0;
                const V3 cellNormal = interfaceN[c];
                const auto& stc = m->stencil[c];
                for (size_t j = 0; j < stc.size(); ++j) {
                    const V3 normal = interfaceN[stc[j]];
                    if (mag(normal) >= TSMALL_ && j != 0) {
                        const V3 n = normal / mag(normal);
                        const double cosAngle = std::max(std::min(dot(cellNormal, n), 1.0), -1.0);
                        avgDiffNormal += std::acos(cosAngle) * mag(normal);
                        weight += mag(normal);
                    }
                }
                if (weight != 0.0) avgDiffNormal /= weight;
                else avgDiffNormal = 0.0;
                const V3 newCellNormal = -normaliseInPlace(g[i], SMALL);
                normalRes[i] = 1.0 - dot(cellNormal, newCellNormal);
                avgAngle[i] = avgDiffNormal;
            }
            int resCounter = 0;
            double avgRes = 0.0, avgNormRes = 0.0;
            for (int i = 0; i < nMixed; ++i) {
                const int c = mixedCells[i];
                const double avgA = avgAngle[i];
                if (avgA > 0.26 && iter > 0) {
                    tooCoarse[c] = 1;
                } else {
                    avgRes += normalRes[i];
                    const double discreteError = 0.01 * (avgA * avgA);
                    const double normRes = discreteError != 0.0 ? normalRes[i] / std::max(discreteError, rdfTol)
                                                                : normalRes[i] / rdfTol;
                    avgNormRes += normRes;
                    resCounter++;
                }
            }
            if (resCounter == 0) {
                resCounter = 1;
                avgRes = 0.0;
                avgNormRes = 0.0;
            }
            if (((avgNormRes / resCounter < rdfRelTol || avgRes / resCounter < rdfTol) && iter >= 1) ||
                iter + 1 == rdfIterations)
                break;
        }
    }

    // reconstruction.C:680-722
    void reconstruct() {
        const auto t0 = std::chrono::steady_clock::now();
        initialize();
        nCuts.assign(mixedCells.size(), 0);
        nCutPts.assign(mixedCells.size(), 0);
        faceClass.assign(mixedCells.size(), 0);
        if (!mixedCells.empty()) {
            if (orientation == 1) calcInterfaceNFromRegAlphaGrad();
            else if (orientation == 2) calcInterfaceNFromIsoRDF();
            else calcInterfaceNFromIsoAlphaGrad();
            const int nMixed = (int)mixedCells.size();
#pragma omp parallel num_threads(nThreads)
            {
                CutCell cc;
                cc.m = m;
                cc.cutFace.m = m;
#pragma omp for schedule(dynamic, 32)
                for (int i = 0; i < nMixed; ++i) {
                    const int c = mixedCells[i];
                    double d = interfaceD[c];
                    V3 C = interfaceC[c], S = interfaceS[c];
                    int nc = 0;
                    cellStatus[i] =
                        cc.findSignedDistance(c, alpha[c], interfaceN[c], splitWarpedFace, d, C, S, &nc);
                    interfaceD[c] = d;
                    interfaceC[c] = C;
                    interfaceS[c] = S;
                    nCuts[i] = nc;
                    nCutPts[i] = cc.nCutPoints;
                    faceClass[i] = cc.faceClass;
                }
            }
        }
        reconstructionTime +=
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // volPointInterpolation value at a point (internal: over pointCells;
    // boundary: over boundary faces with the boundary U values)
    V3 pointU(int p) const {
        V3 t = VZERO;
        if (!m->isPatchPoint[p]) {
            const auto& pc = m->pointCells[p];
            for (size_t i = 0; i < pc.size(); ++i) t += m->pointW[p][i] * U[pc[i]];
        } else {
            const auto& pf = m->bPointFaces[p];
            for (size_t i = 0; i < pf.size(); ++i) t += m->bPointW[p][i] * Ub[pf[i] - m->nIF];
        }
        return t;
    }

    // interpolationCellPoint<vector>::interpolate(position, celli): find the
    // tet (cell centre, face base point 0, fan triangle) holding the position
    // and interpolate barycentrically from the cell value and point values.
    V3 interpolateU(V3 pos, int celli) const {
        const double tol = SMALL;
        const V3 cc = m->C[celli];
        int bestFace = -1, bestTri = -1;
        double bestMin = -1e300;
        double bw[4] = {0, 0, 0, 0};
        int btri[3] = {0, 0, 0};
        for (int f : m->cells[celli]) {
            const int n = m->faceSize(f);
            const int* fp = &m->fPts[m->fOff[f]];
            for (int t = 1; t < n - 1; ++t) {
                int p0 = fp[0], p1 = fp[t], p2 = fp[t + 1];
                if (m->own[f] != celli) std::swap(p1, p2);
                // tet (a=cc, b=p0, c=p1, d=p2); barycentric via e_i = x_i - d
                const V3 a = cc, b = m->points[p0], c = m->points[p1], d = m->points[p2];
                const V3 e0 = a - d, e1 = b - d, e2 = c - d;
                const double detT = dot(e0, cross(e1, e2));
                double w[4];
                if (std::fabs(detT) < SMALL) {
                    w[0] = w[1] = w[2] = w[3] = -1.0;
                } else {
                    const V3 r = pos - d;
                    w[0] = dot(r, cross(e1, e2)) / detT;
                    w[1] = dot(e0, cross(r, e2)) / detT;
                    w[2] = dot(e0, cross(e1, r)) / detT;
                    w[3] = 1.0 - w[0] - w[1] - w[2];
                }
                const double mn = std::min(std::min(w[0], w[1]), std::min(w[2], w[3]));
                if (mn + tol > 0.0 && g_mutation != 2) {
                    return w[0] * U[celli] + w[1] * pointU(p0) + w[2] * pointU(p1) + w[3] * pointU(p2);
                }
                if (g_mutation == 2) {
                    // planted mistake: the tet of the smallest minimum weight
                    if (bestMin == -1e300 || -mn > bestMin) {
                        bestMin = -mn;
                        for (int q = 0; q < 4; ++q) bw[q] = w[q];
                        btri[0] = p0; btri[1] = p1; btri[2] = p2;
                    }
                    continue;
                }
                if (mn > bestMin) {
                    bestMin = mn;
                    bestFace = f;
                    bestTri = t;
                    for (int q = 0; q < 4; ++q) bw[q] = w[q];
                    btri[0] = p0; btri[1] = p1; btri[2] = p2;
                }
            }
        }
        (void)bestFace; (void)bestTri;
        return bw[0] * U[celli] + bw[1] * pointU(btri[0]) + bw[2] * pointU(btri[1]) +
               bw[3] * pointU(btri[2]);
    }

    double faceValue(const std::vector<double>& f, int facei) const {
        if (m->isInternal(facei)) return f[facei];
        if (m->boundaryEmpty(facei)) return 0.0;
        return f[facei];
    }

    // advection.C:85-221
    void timeIntegratedFlux(double dt) {
        const int nMixed = (int)mixedCells.size();
        Un0.assign(nMixed, 0.0);
#pragma omp parallel num_threads(nThreads)
        {
            CutFace cf;
            cf.m = m;
#pragma omp for schedule(dynamic, 32)
            for (int i = 0; i < nMixed; ++i) {
                if (cellStatus[i] != 0) continue;
                const int cellI = mixedCells[i];
                const V3 normalI = interfaceN[cellI];
                const double distanceI = interfaceD[cellI];
                const V3 centreI = interfaceC[cellI];
                const double un0 = dot(interpolateU(centreI, cellI), normalI);
                Un0[i] = un0;
                for (int facei : m->cells[cellI]) {
                    if (m->isInternal(facei)) {
                        bool isDownwindFace = false;
                        if (cellI == m->own[facei]) {
                            if (phi[facei] >= 0.0) isDownwindFace = true;
                        } else {
                            if (phi[facei] < 0.0) isDownwindFace = true;
                        }
                        if (isDownwindFace)
                            dVf[facei] = cf.timeIntegratedFaceFlux(facei, normalI, distanceI, un0, dt,
                                                                   phi[facei], m->magSf[facei]);
                    } else {
                        // boundary surface faces (advection.C:167-217)
                        if (m->boundaryEmpty(facei)) continue;
                        const double phiP = phi[facei];
                        if (phiP >= 0.0)
                            dVf[facei] = cf.timeIntegratedFaceFlux(facei, normalI, distanceI, un0, dt,
                                                                   phiP, m->magSf[facei]);
                    }
                }
            }
        }
    }

    void setDownwindFaces(int cellI, std::vector<int>& downwindFaces) const {
        downwindFaces.clear();
        for (int facei : m->cells[cellI]) {
            const double ph = faceValue(phi, facei);
            if (m->own[facei] == cellI) {
                if (ph >= 0) downwindFaces.push_back(facei);
            } else if (ph < 0) {
                downwindFaces.push_back(facei);
            }
        }
    }

    double netFlux(int cellI, const std::vector<double>& f) const {
        double dV = 0.0;
        for (int facei : m->cells[cellI]) {
            const double v = faceValue(f, facei);
            if (m->own[facei] == cellI) dV += v;
            else dV -= v;
        }
        return dV;
    }

    // advectionTemplates.C:218-349
    void boundFlux(const std::vector<char>& nextToInterface, std::vector<double>& corr,
                   std::vector<int>& correctedFaces, double dt, const double* Sp, const double* Su) {
        const double aTol = 100.0 * SMALL;
        const double rDeltaT = 1.0 / dt;
        correctedFaces.clear();
        std::vector<int> downwindFaces, facesToPass;
        std::vector<double> dVfmax, phiL;
        // schedule statistics (test diagnostics): candidates, cells that write
        // corrections, and the depth of the write-dependency chain
        std::vector<int> depth(m->nC, 0);
        int nCandStat = 0, nWriters = 0, maxDepth = 0;
        for (int cellk = 0; cellk < m->nC; ++cellk) {
            const int celli = g_mutation == 3 ? m->nC - 1 - cellk : cellk;
            if (!nextToInterface[celli]) continue;
            if (!(alpha[celli] < -aTol || alpha[celli] > 1.0 + aTol)) continue;
            nCandStat++;
            bool wrote = false;
            const double Vi = m->V[celli];
            double alphaOvershoot = pos0(alpha[celli] - 1.0) * (alpha[celli] - 1.0) +
                                    neg0(alpha[celli]) * alpha[celli];
            double fluidToPassOn = alphaOvershoot * Vi;
            int nFacesToPassFluidThrough = 1;
            bool firstLoop = true;
            for (int iter = 0; iter < 10; iter++) {
                if (std::fabs(alphaOvershoot) < aTol || nFacesToPassFluidThrough == 0) break;
                facesToPass.clear();
                dVfmax.clear();
                phiL.clear();
                setDownwindFaces(celli, downwindFaces);
                double dVftot = 0.0;
                nFacesToPassFluidThrough = 0;
                for (int facei : downwindFaces) {
                    const double phif = faceValue(phi, facei);
                    const double dVff = faceValue(dVf, facei) + faceValue(corr, facei);
                    const double maxExtraFaceFluidTrans =
                        std::fabs(pos0(fluidToPassOn) * phif * dt - dVff);
                    if (maxExtraFaceFluidTrans / Vi > aTol) {
                        facesToPass.push_back(facei);
                        phiL.push_back(phif);
                        dVfmax.push_back(maxExtraFaceFluidTrans);
                        dVftot += std::fabs(phif * dt);
                    }
                }
                for (size_t fi = 0; fi < facesToPass.size(); ++fi) {
                    const int facei = facesToPass[fi];
                    double fluidToPassThroughFace =
                        std::fabs(fluidToPassOn) * std::fabs(phiL[fi] * dt) / dVftot;
                    nFacesToPassFluidThrough += (int)pos0(dVfmax[fi] - fluidToPassThroughFace);
                    fluidToPassThroughFace = std::min(fluidToPassThroughFace, dVfmax[fi]);
                    double dVff = faceValue(corr, facei);
                    dVff += sgn(phiL[fi]) * sgn(fluidToPassOn) * fluidToPassThroughFace;
                    if (m->isInternal(facei) || !m->boundaryEmpty(facei)) corr[facei] = dVff;
                    wrote = true;
                    if (firstLoop) correctedFaces.push_back(facei);
                }
                firstLoop = false;
                double num = alphaOld[celli] * rDeltaT;
                if (Su) num = num + Su[celli];
                num = num - netFlux(celli, dVf) / Vi * rDeltaT - netFlux(celli, corr) / Vi * rDeltaT;
                const double den = Sp ? (rDeltaT - Sp[celli]) : rDeltaT;
                const double alpha1New = num / den;
                alphaOvershoot = pos0(alpha1New - 1.0) * (alpha1New - 1.0) + neg0(alpha1New) * alpha1New;
                fluidToPassOn = alphaOvershoot * Vi;
            }
            if (wrote) {
                // longest chain of writers j < celli upwind of celli
                int dmax = 0;
                for (int f : m->cells[celli]) {
                    if (!m->isInternal(f)) continue;
                    const int o = m->own[f] == celli ? m->nei[f] : m->own[f];
                    const bool upwindOfCell = (m->own[f] == celli) ? (phi[f] < 0) : (phi[f] >= 0);
                    if (o < celli && upwindOfCell) dmax = std::max(dmax, depth[o]);
                }
                depth[celli] = dmax + 1;
                nWriters++;
                maxDepth = std::max(maxDepth, depth[celli]);
            }
        }
        schedStats.push_back(nCandStat);
        schedStats.push_back(nWriters);
        schedStats.push_back(maxDepth);
    }
    std::vector<int> schedStats;

    // advection.C:54-82
    void extendMarkedCells(std::vector<char>& marked) const {
        std::vector<char> markedFace(m->nF, 0);
        for (int c = 0; c < m->nC; ++c)
            if (marked[c])
                for (int f : m->cells[c]) markedFace[f] = 1;
        for (int f = 0; f < m->nIF; ++f)
            if (markedFace[f]) { marked[m->own[f]] = 1; marked[m->nei[f]] = 1; }
        for (int f = m->nIF; f < m->nF; ++f)
            if (markedFace[f]) marked[m->own[f]] = 1;
    }

    std::vector<double> boundLog;   // per pass: min, max before the pass

    // advectionTemplates.C:118-215
    void limitFlux(double dt, const double* Sp, const double* Su) {
        const double aTol = 100.0 * SMALL;
        auto gMax = [&]() { double r = -1e300; for (double a : alpha) r = std::max(r, a); return r; };
        auto gMin = [&]() { double r = 1e300; for (double a : alpha) r = std::min(r, a); return r; };
        double maxAlphaMinus1 = gMax() - 1.0;
        double minAlpha = gMin();
        boundLog.clear();
        schedStats.clear();
        std::vector<double> corr(m->nF, 0.0);
        std::vector<char> needBounding(m->nC, 0);
        for (int c : mixedCells) needBounding[c] = 1;
        extendMarkedCells(needBounding);
        std::vector<int> correctedFaces;
        for (int n = 0; n < nAlphaBounds; n++) {
            if (maxAlphaMinus1 > aTol || minAlpha < -aTol) {
                boundLog.push_back(minAlpha);
                boundLog.push_back(maxAlphaMinus1);
                std::fill(corr.begin(), corr.end(), 0.0);
                boundFlux(needBounding, corr, correctedFaces, dt, Sp, Su);
                std::vector<char> alreadyUpdated(m->nF, 0);
                for (int facei : correctedFaces) {
                    if (alreadyUpdated[facei]) continue;
                    alreadyUpdated[facei] = 1;
                    const int o = m->own[facei];
                    alpha[o] -= faceValue(corr, facei) / m->V[o];
                    if (m->isInternal(facei)) {
                        const int nb = m->nei[facei];
                        alpha[nb] += faceValue(corr, facei) / m->V[nb];
                    }
                    const double corrVf = faceValue(dVf, facei) + faceValue(corr, facei);
                    if (m->isInternal(facei) || !m->boundaryEmpty(facei)) dVf[facei] = corrVf;
                }
            } else {
                break;
            }
            maxAlphaMinus1 = gMax() - 1.0;
            minAlpha = gMin();
        }
    }

    // advectionTemplates.C:352-418 (static mesh: no motion, no overset)
    void advect(double dt, const double* Sp, const double* Su) {
        const auto t0 = std::chrono::steady_clock::now();
        const double rDeltaT = 1.0 / dt;
        alphaOld = alpha;
        // dVf = upwind(phi).flux(alpha) * dt, lambda-form face interpolation
        for (int f = 0; f < m->nIF; ++f) {
            const double lambda = pos0(phi[f]);
            const double af = lambda * (alpha[m->own[f]] - alpha[m->nei[f]]) + alpha[m->nei[f]];
            dVf[f] = (phi[f] * af) * dt;
        }
        for (int f = m->nIF; f < m->nF; ++f) {
            if (m->boundaryEmpty(f)) { dVf[f] = 0.0; continue; }
            dVf[f] = (phi[f] * alpha[m->own[f]]) * dt;   // zeroGradient boundary value
        }
        timeIntegratedFlux(dt);
        // fvc::surfaceIntegrate(dVf)
        std::vector<double> sI(m->nC, 0.0);
        for (int f = 0; f < m->nIF; ++f) {
            sI[m->own[f]] += dVf[f];
            sI[m->nei[f]] -= dVf[f];
        }
        for (int f = m->nIF; f < m->nF; ++f) {
            if (m->boundaryEmpty(f)) continue;
            sI[m->own[f]] += dVf[f];
        }
        for (int c = 0; c < m->nC; ++c) {
            sI[c] = sI[c] / m->V[c];
            double num = alphaOld[c] * rDeltaT;
            if (Su) num = num + Su[c];
            num = num - sI[c] * rDeltaT;
            const double den = Sp ? (rDeltaT - Sp[c]) : rDeltaT;
            alpha[c] = num / den;
        }
        limitFlux(dt, Sp, Su);
        // advection.C:291-308
        if (snapTol > 0.0) {
            for (double& a : alpha)
                a = a * pos0(a - snapTol) * neg0(a - (1.0 - snapTol)) + pos0(a - (1.0 - snapTol));
        }
        if (clip) {
            for (double& a : alpha) a = std::min(1.0, std::max(0.0, a));
        }
        advectionTime += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int f = 0; f < m->nF; ++f) alphaPhi[f] = dVf[f] / dt;
    }
};

}  // namespace oracle

using namespace oracle;

extern "C" {

void* or_mesh_new(int nP, const double* pts, int nF, int nIF, const int* fOff, const int* fPts,
                  const int* own, const int* nei, int nC, int nPatch, const int* pStart,
                  const int* pSize, const int* pType) {
    Mesh* m = new Mesh();
    m->nP = nP; m->nF = nF; m->nIF = nIF; m->nC = nC;
    m->points.resize(nP);
    for (int p = 0; p < nP; ++p) m->points[p] = v3(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]);
    m->fOff.assign(fOff, fOff + nF + 1);
    m->fPts.assign(fPts, fPts + fOff[nF]);
    m->own.assign(own, own + nF);
    m->nei.assign(nei, nei + nIF);
    m->patchStart.assign(pStart, pStart + nPatch);
    m->patchSize.assign(pSize, pSize + nPatch);
    m->patchType.assign(pType, pType + nPatch);
    buildMesh(*m);
    return m;
}

void or_mesh_free(void* h) { delete (Mesh*)h; }

void or_mesh_geometry(void* h, double* Cf, double* Sf, double* magSf, double* C, double* V,
                      double* flatness) {
    Mesh* m = (Mesh*)h;
    for (int f = 0; f < m->nF; ++f) {
        if (Cf) { Cf[3 * f] = m->Cf[f].x; Cf[3 * f + 1] = m->Cf[f].y; Cf[3 * f + 2] = m->Cf[f].z; }
        if (Sf) { Sf[3 * f] = m->Sf[f].x; Sf[3 * f + 1] = m->Sf[f].y; Sf[3 * f + 2] = m->Sf[f].z; }
        if (magSf) magSf[f] = m->magSf[f];
        if (flatness) flatness[f] = m->flatness[f];
    }
    for (int c = 0; c < m->nC; ++c) {
        if (C) { C[3 * c] = m->C[c].x; C[3 * c + 1] = m->C[c].y; C[3 * c + 2] = m->C[c].z; }
        if (V) V[c] = m->V[c];
    }
}

int or_mesh_stencil_size(void* h, int c) { return (int)((Mesh*)h)->stencil[c].size(); }

static std::vector<V3> toPts(int n, const double* p) {
    std::vector<V3> r(n);
    for (int i = 0; i < n; ++i) r[i] = v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    return r;
}

int or_calc_sub_face(int n, const double* pts, const double* nrm, double d, double* area,
                     double* centre, int* nSub, int* nCut) {
    CutFace cf;
    const int st = cf.calcSubFace(toPts(n, pts), v3(nrm[0], nrm[1], nrm[2]), d);
    area[0] = cf.subFaceArea.x; area[1] = cf.subFaceArea.y; area[2] = cf.subFaceArea.z;
    centre[0] = cf.subFaceCentre.x; centre[1] = cf.subFaceCentre.y; centre[2] = cf.subFaceCentre.z;
    if (nSub) *nSub = (int)cf.subFacePoints.size();
    if (nCut) *nCut = (int)cf.interfacePoints.size();
    return st;
}

double or_time_integrated_area(int n, const double* pts, const double* nrm, double d,
                               const double* pTimes, double Un0, double dt, double magSf) {
    CutFace cf;
    std::vector<double> t(pTimes, pTimes + n);
    return cf.timeIntegratedArea(toPts(n, pts), v3(nrm[0], nrm[1], nrm[2]), d, t, Un0, dt, magSf);
}

double or_time_integrated_face_flux(void* h, int face, const double* nrm, double d, double Un0,
                                    double dt, double phi, double magSf) {
    CutFace cf;
    cf.m = (Mesh*)h;
    return cf.timeIntegratedFaceFlux(face, v3(nrm[0], nrm[1], nrm[2]), d, Un0, dt, phi, magSf);
}

int or_calc_sub_cell(void* h, int cell, const double* nrm, double d, int split, double* out) {
    CutCell cc;
    cc.m = (Mesh*)h;
    cc.cutFace.m = cc.m;
    if (split) cc.getLocalPointFieldAndFaceList(cell);
    const int st = cc.calcSubCell(cell, v3(nrm[0], nrm[1], nrm[2]), d, split != 0);
    // out: VOF, subCellVolume, subCellCentre(3), interfaceCentre(3), interfaceArea(3)
    out[0] = cc.VOF; out[1] = cc.subCellVolume;
    out[2] = cc.subCellCentre.x; out[3] = cc.subCellCentre.y; out[4] = cc.subCellCentre.z;
    out[5] = cc.interfaceCentre.x; out[6] = cc.interfaceCentre.y; out[7] = cc.interfaceCentre.z;
    out[8] = cc.interfaceArea.x; out[9] = cc.interfaceArea.y; out[10] = cc.interfaceArea.z;
    return st;
}

int or_find_signed_distance(void* h, int cell, double alpha, const double* nrm, int split,
                            double* out) {
    CutCell cc;
    cc.m = (Mesh*)h;
    cc.cutFace.m = cc.m;
    double d = 0.0;
    V3 C = VZERO, S = VZERO;
    int nc = 0;
    const int st = cc.findSignedDistance(cell, alpha, v3(nrm[0], nrm[1], nrm[2]), split != 0, d, C, S, &nc);
    out[0] = d; out[1] = C.x; out[2] = C.y; out[3] = C.z; out[4] = S.x; out[5] = S.y; out[6] = S.z;
    out[7] = nc;
    out[8] = cc.nCutPoints;
    out[9] = (double)cc.faceClass;
    return st;
}

void* or_solver_new(void* mesh, double mixedCellTol, int nAlphaBounds, double snapTol, int clip,
                    int splitWarpedFace, int nThreads) {
    Solver* s = new Solver();
    s->m = (Mesh*)mesh;
    s->mixedCellTol = mixedCellTol;
    s->nAlphaBounds = nAlphaBounds;
    s->snapTol = snapTol;
    s->clip = clip != 0;
    s->splitWarpedFace = splitWarpedFace != 0;
    s->nThreads = nThreads < 1 ? 1 : nThreads;
    s->init();
    return s;
}

void or_solver_free(void* h) { delete (Solver*)h; }

// 0: isoAlphaGrad / LS, 1: alphaGrad, 2: isoRDF / RDF
void or_solver_set_orientation(void* h, int method) { ((Solver*)h)->orientation = method; }
// gradient scheme of the alphaGrad normal: 0 Gauss linear, 1 Gauss pointLinear
void or_solver_set_grad_scheme(void* h, int scheme) { ((Solver*)h)->gradScheme = scheme; }
// the alphaGrad normal of every cell from the given alpha (tests)
void or_grad_normals(void* h, const double* a, double* out) {
    Solver* s = (Solver*)h;
    s->alpha.assign(a, a + s->m->nC);
    s->interfaceN.assign(s->m->nC, VZERO);
    s->calcInterfaceNFromRegAlphaGrad();
    for (int c = 0; c < s->m->nC; ++c) {
        out[3 * c] = s->interfaceN[c].x;
        out[3 * c + 1] = s->interfaceN[c].y;
        out[3 * c + 2] = s->interfaceN[c].z;
    }
}
// isoRDF keys tol, relTol, iterations (reconstruction.C:513-515)
void or_solver_set_rdf(void* h, double tol, double relTol, int iterations) {
    Solver* s = (Solver*)h;
    s->rdfTol = tol;
    s->rdfRelTol = relTol;
    s->rdfIterations = iterations;
}
int or_rdf_iterations_run(void* h) { return ((Solver*)h)->rdfIterationsRun; }

void or_set_alpha(void* h, const double* a) {
    Solver* s = (Solver*)h;
    s->alpha.assign(a, a + s->m->nC);
}

void or_set_flow(void* h, const double* phi, const double* U, const double* Ub) {
    Solver* s = (Solver*)h;
    s->phi.assign(phi, phi + s->m->nF);
    for (int c = 0; c < s->m->nC; ++c) s->U[c] = v3(U[3 * c], U[3 * c + 1], U[3 * c + 2]);
    const int nb = s->m->nF - s->m->nIF;
    for (int b = 0; b < nb; ++b)
        s->Ub[b] = Ub ? v3(Ub[3 * b], Ub[3 * b + 1], Ub[3 * b + 2]) : VZERO;
}

void or_reconstruct(void* h) { ((Solver*)h)->reconstruct(); }

void or_advect(void* h, double dt, const double* Sp, const double* Su) {
    ((Solver*)h)->advect(dt, Sp, Su);
}

void or_get_alpha(void* h, double* a) {
    Solver* s = (Solver*)h;
    std::memcpy(a, s->alpha.data(), sizeof(double) * s->m->nC);
}
void or_get_dvf(void* h, double* a) {
    Solver* s = (Solver*)h;
    std::memcpy(a, s->dVf.data(), sizeof(double) * s->m->nF);
}
void or_get_alpha_phi(void* h, double* a) {
    Solver* s = (Solver*)h;
    std::memcpy(a, s->alphaPhi.data(), sizeof(double) * s->m->nF);
}
// advection::getRhoPhi with constant densities (advection.H:329-343).
void or_get_rho_phi(void* h, double rho1, double rho2, double* out) {
    Solver* s = (Solver*)h;
    const double drho = rho1 - rho2;
    for (int f = 0; f < s->m->nF; ++f) out[f] = drho * s->alphaPhi[f] + rho2 * s->phi[f];
}

// alphaEqnSubCycle.H:30-60 with alphaEqn.H:9-20.  Time::subCycle divides
// deltaT by the number of sub-cycles; every sub-cycle adds
// (subDeltaT/totalDeltaT) rhoPhi to a sum that starts at zero.
void or_advect_sub_cycles(void* h, double dt, int nSub, const double* Sp, const double* Su, double rho1,
                          double rho2, double* rhoPhi) {
    Solver* s = (Solver*)h;
    const int nF = s->m->nF;
    if (nSub <= 1) {
        s->reconstruct();
        s->advect(dt, Sp, Su);
        or_get_rho_phi(h, rho1, rho2, rhoPhi);
        return;
    }
    const double subDt = dt / nSub;
    const double w = subDt / dt;
    std::vector<double> sum(nF, 0.0), rp(nF);
    for (int k = 0; k < nSub; ++k) {
        s->reconstruct();
        s->advect(subDt, Sp, Su);
        or_get_rho_phi(h, rho1, rho2, rp.data());
        for (int f = 0; f < nF; ++f) sum[f] += w * rp[f];
    }
    std::memcpy(rhoPhi, sum.data(), sizeof(double) * nF);
}

// reconstruction::interface (reconstruction.C:787-835): every mixed cell of
// the last reconstruct is cut again by its plane (n, d) through its mesh faces
// (no warped-face split); one polygon of interfacePoints per cell the plane
// cuts, cells ascending.  With pts == NULL only the sizes are returned.
void or_plic_surface(void* h, int* nFaces, int* nPts, int* cells, int* offsets, double* pts) {
    Solver* s = (Solver*)h;
    CutCell cc;
    cc.m = s->m;
    int nf = 0, np = 0;
    for (int c : s->mixedCells) {
        if (cc.calcSubCell(c, s->interfaceN[c], s->interfaceD[c], false) != 0) continue;
        const std::vector<V3> poly = cc.interfacePoints();
        if (pts) {
            cells[nf] = c;
            offsets[nf] = np;
            for (size_t k = 0; k < poly.size(); ++k) {
                pts[3 * (np + k)] = poly[k].x;
                pts[3 * (np + k) + 1] = poly[k].y;
                pts[3 * (np + k) + 2] = poly[k].z;
            }
        }
        nf++;
        np += (int)poly.size();
    }
    if (pts) offsets[nf] = np;
    *nFaces = nf;
    *nPts = np;
}

// reconstruction::mapAlphaField (reconstruction.C:725-784) with per-cell
// n (3 per cell) and d: cells with lower <= alpha <= upper take the volume
// fraction of their plane cut through the mesh faces.
void or_map_alpha_field(void* h, const double* n, const double* d, double lower, double upper) {
    Solver* s = (Solver*)h;
    CutCell cc;
    cc.m = s->m;
    cc.cutFace.m = s->m;
    for (int c = 0; c < s->m->nC; ++c) {
        if (s->alpha[c] >= lower && s->alpha[c] <= upper) {
            cc.calcSubCell(c, v3(n[3 * c], n[3 * c + 1], n[3 * c + 2]), d[c], false);
            s->alpha[c] = cc.VOF;
        }
    }
}

int or_n_mixed(void* h) { return (int)((Solver*)h)->mixedCells.size(); }
void or_get_cut_class(void* h, int* nCutPts, uint64_t* faceClass) {
    Solver* s = (Solver*)h;
    for (size_t i = 0; i < s->mixedCells.size(); ++i) {
        nCutPts[i] = s->nCutPts[i];
        faceClass[i] = s->faceClass[i];
    }
}
// per mixed cell: cell, status, nCuts, n(3), d, C(3), S(3), Un0
void or_get_mixed(void* h, int* cells, int* status, int* ncuts, double* n, double* d, double* C,
                  double* S, double* un0) {
    Solver* s = (Solver*)h;
    for (size_t i = 0; i < s->mixedCells.size(); ++i) {
        const int c = s->mixedCells[i];
        if (cells) cells[i] = c;
        if (status) status[i] = s->cellStatus[i];
        if (ncuts) ncuts[i] = i < s->nCuts.size() ? s->nCuts[i] : 0;
        if (n) { n[3 * i] = s->interfaceN[c].x; n[3 * i + 1] = s->interfaceN[c].y; n[3 * i + 2] = s->interfaceN[c].z; }
        if (d) d[i] = s->interfaceD[c];
        if (C) { C[3 * i] = s->interfaceC[c].x; C[3 * i + 1] = s->interfaceC[c].y; C[3 * i + 2] = s->interfaceC[c].z; }
        if (S) { S[3 * i] = s->interfaceS[c].x; S[3 * i + 1] = s->interfaceS[c].y; S[3 * i + 2] = s->interfaceS[c].z; }
        if (un0) un0[i] = i < s->Un0.size() ? s->Un0[i] : 0.0;
    }
}
// the interface normals before plane placement are exposed through or_get_mixed
void or_set_normals(void* h, const double* n) {
    // override the LS normals of the current mixed cells (used by parity tests
    // that feed identical normals to both sides)
    Solver* s = (Solver*)h;
    for (size_t i = 0; i < s->mixedCells.size(); ++i)
        s->interfaceN[s->mixedCells[i]] = v3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
}
void or_timers(void* h, double* out) {
    Solver* s = (Solver*)h;
    out[0] = s->reconstructionTime;
    out[1] = s->advectionTime;
}
int or_bound_log(void* h, double* out, int cap) {
    Solver* s = (Solver*)h;
    const int n = std::min(cap, (int)s->boundLog.size());
    for (int i = 0; i < n; ++i) out[i] = s->boundLog[i];
    return (int)s->boundLog.size();
}

int or_sched_stats(void* h, int* out, int cap) {
    Solver* s = (Solver*)h;
    const int n = std::min(cap, (int)s->schedStats.size());
    for (int i = 0; i < n; ++i) out[i] = s->schedStats[i];
    return (int)s->schedStats.size();
}

// Point value of U by volPointInterpolation (exposed for tests)
void or_point_u(void* h, int p, double* out) {
    Solver* s = (Solver*)h;
    const V3 u = s->pointU(p);
    out[0] = u.x; out[1] = u.y; out[2] = u.z;
}
void or_interpolate_u(void* h, const double* pos, int cell, double* out) {
    Solver* s = (Solver*)h;
    const V3 u = s->interpolateU(v3(pos[0], pos[1], pos[2]), cell);
    out[0] = u.x; out[1] = u.y; out[2] = u.z;
}

void or_set_mutation(int k) { g_mutation = k; }

// limitFlux (advectionTemplates.C:118-215) alone on a given state (exposed for
// tests): alpha after the update, alpha.oldTime(), dVf and the mixed cells;
// phi comes from or_set_flow.  Writes the bounded alpha and dVf.
void or_limit_flux(void* h, double dt, const double* alpha, const double* alphaOld, const double* dVf,
                   const int* mixed, int nMixed, double* alphaOut, double* dVfOut) {
    Solver* s = (Solver*)h;
    s->alpha.assign(alpha, alpha + s->m->nC);
    s->alphaOld.assign(alphaOld, alphaOld + s->m->nC);
    s->dVf.assign(dVf, dVf + s->m->nF);
    s->mixedCells.assign(mixed, mixed + nMixed);
    s->limitFlux(dt, nullptr, nullptr);
    std::copy(s->alpha.begin(), s->alpha.end(), alphaOut);
    std::copy(s->dVf.begin(), s->dVf.end(), dVfOut);
}

// LS gradient of arbitrary stencil data (exposed for tests)
void or_ls_grad(int n, const double* pos, const double* val, double* out) {
    std::vector<V3> p = toPts(n, pos);
    std::vector<double> v(val, val + n);
    const V3 g = lsGrad(p, v);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

// ---------------------------------------------------------------------------
// Harness restatement (applications/test/plicVofAdvectionFoam)
// ---------------------------------------------------------------------------
// updateU.H:66-91 at cell centres and face centres (unscaled), glibc sin.
void or_deformation_field(void* h, double* Uc, double* Ub, double* phi) {
    Mesh* m = (Mesh*)h;
    const double PI = M_PI;
    auto U = [&](V3 p) {
        return v3(2 * ((std::sin(PI * p.x)) * (std::sin(PI * p.x))) * std::sin(2 * PI * p.y) * std::sin(2 * PI * p.z),
                  -std::sin(2 * PI * p.x) * ((std::sin(PI * p.y)) * (std::sin(PI * p.y))) * std::sin(2 * PI * p.z),
                  -std::sin(2 * PI * p.x) * std::sin(2 * PI * p.y) * ((std::sin(PI * p.z)) * (std::sin(PI * p.z))));
    };
    for (int c = 0; c < m->nC; ++c) {
        const V3 u = U(m->C[c]);
        Uc[3 * c] = u.x; Uc[3 * c + 1] = u.y; Uc[3 * c + 2] = u.z;
    }
    for (int f = 0; f < m->nF; ++f) phi[f] = dot(U(m->Cf[f]), m->Sf[f]);
    // U boundary condition: fixedValue (0 0 0) on every wall
    if (Ub)
        for (int b = 0; b < m->nF - m->nIF; ++b) Ub[3 * b] = Ub[3 * b + 1] = Ub[3 * b + 2] = 0.0;
}

// CourantNo.H / alphaCourantNo.H:95-107: returns max_c(sum|phi|/V) and the
// same restricted to 0.01 <= alpha <= 0.99 (multiply by 0.5*dt outside).
void or_courant(void* h, const double* phi, const double* alpha, double* out) {
    Mesh* m = (Mesh*)h;
    std::vector<double> sumPhi(m->nC, 0.0);
    for (int f = 0; f < m->nIF; ++f) {
        sumPhi[m->own[f]] += std::fabs(phi[f]);
        sumPhi[m->nei[f]] += std::fabs(phi[f]);
    }
    for (int f = m->nIF; f < m->nF; ++f) {
        if (m->boundaryEmpty(f)) continue;
        sumPhi[m->own[f]] += std::fabs(phi[f]);
    }
    double co = -1e300, aco = -1e300;
    for (int c = 0; c < m->nC; ++c) {
        const double r = sumPhi[c] / m->V[c];
        co = std::max(co, r);
        const double ind = pos0(alpha[c] - 0.01) * pos0(0.99 - alpha[c]);
        aco = std::max(aco, (ind * sumPhi[c]) / m->V[c]);
    }
    out[0] = co;
    out[1] = aco;
}

}  // extern "C"
