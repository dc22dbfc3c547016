// Device-side FP64 geometry of the SimPLIC alpha transport (gfx950).
//
// Every routine reproduces the operation order of the reference so that the
// results are bit-identical to an IEEE-754 x86-64 build of the reference:
// the translation unit is compiled with -ffp-contract=off (no FMA fusion) and
// uses the correctly rounded double '/' and sqrt.
//
// Face cuts are evaluated one face at a time into a FaceCut record; a cell cut
// combines the records of its faces in cell-face order.  The records are the
// unit of parallelism of the kernels (one lane per face and cut level) and the
// combination order is what keeps the sums bit-exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svp {

// OpenFOAM double-precision constants
#define SVP_SMALL 1e-15
#define SVP_VSMALL 1e-300
#define SVP_ROOTVSMALL 1e-150
#define SVP_TSMALL 1e-14

// Size limits of the device cut routines (checked on the host at create time).
constexpr int MAX_FACE_POINTS = 8;    // points of one (local) face
constexpr int MAX_IF_POINTS = 4;      // plane crossings recorded per face
constexpr int MAX_CELL_FACES = 24;    // local faces of one cell (split-warped hex: 24)
constexpr int MAX_CELL_POINTS = 24;   // local points of one cell
// cells without warped-face splitting (mesh faces as they are)
constexpr int NS_CELL_FACES = 8;
constexpr int NS_CELL_POINTS = 16;
constexpr int NS_FACE_IDX = 48;
// cells with warped faces split into triangles around the face centre
constexpr int SP_FACE_IDX = 96;
// general polyhedra (no split) beyond the fast path: 2:1-refined hexahedra
// (up to 24 faces, 20 points) and agglomerated or polyhedral cells
constexpr int GEN_CELL_FACES = 24;
constexpr int GEN_CELL_POINTS = 32;
constexpr int GEN_FACE_IDX = GEN_CELL_FACES * MAX_FACE_POINTS;

struct V3 {
    double x, y, z;
};
__host__ __device__ inline V3 v3(double a, double b, double c) { return V3{a, b, c}; }
// component-wise choice: a conditional V3 assignment written this way stays in
// registers (a struct assignment under a condition can become a store through
// a selected address, which moves the variables to scratch memory)
__host__ __device__ inline V3 vsel(bool c, V3 a, V3 b) {
    return V3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__host__ __device__ inline V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__host__ __device__ inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double magSqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ inline double mag(V3 a) { return sqrt(magSqr(a)); }
__host__ __device__ inline double sgn(double s) { return (s >= 0.0) ? 1.0 : -1.0; }
__host__ __device__ inline double pos0(double s) { return (s >= 0.0) ? 1.0 : 0.0; }
__host__ __device__ inline double neg0(double s) { return (s <= 0.0) ? 1.0 : 0.0; }
__device__ inline V3 ld3(const double* p, int i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ inline void st3(double* p, int i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
constexpr V3 VZERO = {0.0, 0.0, 0.0};

// face::centre of an ordered point list (OpenFOAM face.C)
__device__ inline V3 faceCentre(const V3* p, int n) {
    if (n == 3) return (1.0 / 3.0) * (p[0] + p[1] + p[2]);
    V3 c = VZERO;
    for (int i = 0; i < n; ++i) c = c + p[i];
    c = c / (double)n;
    double sumA = 0.0;
    V3 sumAc = VZERO;
    for (int i = 0; i < n; ++i) {
        const V3 nx = p[(i + 1) % n];
        const V3 ttc = p[i] + nx + c;
        const double ta = mag(cross(p[i] - c, nx - c));
        sumA += ta;
        sumAc = sumAc + ta * ttc;
    }
    if (sumA > SVP_VSMALL) return sumAc / (3.0 * sumA);
    return c;
}

__device__ inline V3 triAreaNormal(V3 a, V3 b, V3 c) { return 0.5 * cross(b - a, c - a); }

// face::areaNormal (OpenFOAM face.C)
__device__ inline V3 faceAreaNormal(const V3* p, int n) {
    if (n == 3) return triAreaNormal(p[0], p[1], p[2]);
    V3 c = VZERO;
    for (int i = 0; i < n; ++i) c = c + p[i];
    c = c / (double)n;
    V3 nrm = VZERO;
    for (int i = 0; i < n; ++i) {
        const V3 nx = (i < n - 1) ? p[i + 1] : p[0];
        nrm = nrm + triAreaNormal(p[i], nx, c);
    }
    return nrm;
}

// primitiveMeshTools::makeFaceCentresAndAreas for one face
__device__ inline void primitiveFaceGeom(const V3* p, int n, V3& fc, V3& fa) {
    if (n == 3) {
        fc = (1.0 / 3.0) * (p[0] + p[1] + p[2]);
        fa = 0.5 * cross(p[1] - p[0], p[2] - p[0]);
        return;
    }
    V3 sumN = VZERO, sumAc = VZERO;
    double sumA = 0.0;
    V3 c = p[0];
    for (int i = 1; i < n; ++i) c = c + p[i];
    c = c / (double)n;
    for (int i = 0; i < n; ++i) {
        const V3 nx = p[i == n - 1 ? 0 : i + 1];
        const V3 th = p[i];
        const V3 cc = th + nx + c;
        const V3 nn = cross(nx - th, c - th);
        const double a = mag(nn);
        sumN = sumN + nn;
        sumA += a;
        sumAc = sumAc + a * cc;
    }
    if (sumA < SVP_ROOTVSMALL) {
        fc = c;
        fa = VZERO;
    } else {
        fc = (1.0 / 3.0) * sumAc / sumA;
        fa = 0.5 * sumN;
    }
}

// Result of cutting one face by the plane n.x + d = 0 (cutFace::calcSubFace)
struct FaceCut {
    int status;          // -1 submerged, 0 cut, 1 empty
    int nIf;             // number of recorded plane crossings
    V3 centre, area;     // sub-face centre and area vector
    V3 ifPts[MAX_IF_POINTS];
};

// cutFace::calcSubFace (cutFace.C:136-259) + calcSubFaceCentreAndArea
// (cutFace.C:37-96).  The sub-face polygon is walked twice instead of being
// stored: once for the mean point, once for the triangle fan.  Returns false
// when more than MAX_IF_POINTS crossings occur.
__device__ inline bool calcSubFace(const V3* fPts, int n, V3 normal, double distance, FaceCut& out,
                                   bool wantArea = true) {
    double dist[MAX_FACE_POINTS];
    int nSub = 0, first = -1;
    for (int i = 0; i < n; ++i) {
        double di = dot(fPts[i], normal) + distance;
        if (fabs(di) < SVP_TSMALL) di += sgn(di) * SVP_TSMALL;
        dist[i] = di;
        if (di < 0.0) {
            nSub++;
            if (first == -1) first = i;
        }
    }
    out.nIf = 0;
    if (nSub == n) {
        out.status = -1;
        out.centre = faceCentre(fPts, n);
        out.area = faceAreaNormal(fPts, n);
        return true;
    }
    if (nSub == 0) {
        out.status = 1;
        out.centre = VZERO;
        out.area = VZERO;
        return true;
    }
    // pass 1: count points, mean, crossings
    bool ok = true;
    int nPts = 0;
    V3 p0 = VZERO, p1 = VZERO, p2 = VZERO;
    V3 fCentre = VZERO;
    for (int i = first; i < first + n; ++i) {
        const int cur = i % n, nxt = (i + 1) % n;
        if (dist[cur] < 0) {
            const V3 q = fPts[cur];
            if (nPts == 0) fCentre = q; else fCentre = fCentre + q;
            if (nPts == 0) p0 = q; else if (nPts == 1) p1 = q; else if (nPts == 2) p2 = q;
            nPts++;
        }
        if ((dist[cur] * dist[nxt]) < 0) {
            const double w = dist[cur] / (dist[cur] - dist[nxt]);
            const V3 q = fPts[cur] + w * (fPts[nxt] - fPts[cur]);
            if (nPts == 0) fCentre = q; else fCentre = fCentre + q;
            if (nPts == 0) p0 = q; else if (nPts == 1) p1 = q; else if (nPts == 2) p2 = q;
            nPts++;
            if (out.nIf < MAX_IF_POINTS) out.ifPts[out.nIf] = q; else ok = false;
            out.nIf++;
        }
    }
    if (out.nIf > MAX_IF_POINTS) out.nIf = MAX_IF_POINTS;
    if (nPts < 3) {
        out.status = -1;
        out.centre = faceCentre(fPts, n);
        out.area = faceAreaNormal(fPts, n);
        return ok;
    }
    out.status = 0;
    if (!wantArea) return ok;
    if (nPts == 3) {
        out.centre = (1.0 / 3.0) * (p0 + p1 + p2);
        out.area = 0.5 * cross(p1 - p0, p2 - p0);
        return ok;
    }
    fCentre = fCentre / (double)nPts;
    // pass 2: triangle fan (p_k, p_k+1, fCentre) in emission order
    V3 sumN = VZERO, sumAc = VZERO;
    double sumA = 0.0;
    V3 prev = VZERO, firstP = VZERO;
    int k = 0;
    auto tri = [&](V3 a, V3 b) {
        const V3 c = a + b + fCentre;
        const V3 nn = cross(b - a, fCentre - a);
        const double ar = mag(nn);
        sumN = sumN + nn;
        sumA += ar;
        sumAc = sumAc + ar * c;
    };
    for (int i = first; i < first + n; ++i) {
        const int cur = i % n, nxt = (i + 1) % n;
        if (dist[cur] < 0) {
            const V3 q = fPts[cur];
            if (k == 0) firstP = q; else tri(prev, q);
            prev = q;
            k++;
        }
        if ((dist[cur] * dist[nxt]) < 0) {
            const double w = dist[cur] / (dist[cur] - dist[nxt]);
            const V3 q = fPts[cur] + w * (fPts[nxt] - fPts[cur]);
            if (k == 0) firstP = q; else tri(prev, q);
            prev = q;
            k++;
        }
    }
    tri(prev, firstP);
    if (sumA < SVP_ROOTVSMALL) {
        out.centre = fCentre;
        out.area = VZERO;
    } else {
        out.centre = (1.0 / 3.0) * sumAc / sumA;
        out.area = 0.5 * sumN;
    }
    return ok;
}

// Element j of a register array of N values, as a select chain (keeps the
// array in registers where a dynamic index would spill it to scratch).
template <int N, class T>
__device__ inline T pick(const T (&a)[N], int j) {
    T r = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k)
        if (j == k) r = a[k];
    return r;
}

// Rotate register arrays left by r (0 <= r < N): a[i] <- a[(i + r) % N], by
// conditional moves with static indices only (binary stages of the shift),
// so the arrays stay in registers.
template <int N, class T>
__device__ inline void rotate_left(T (&a)[N], int r) {
#pragma unroll
    for (int s = 1; s < N; s <<= 1) {
        if (r & s) {
            T t[N];
#pragma unroll
            for (int i = 0; i < N; ++i) t[i] = a[(i + s) % N];
#pragma unroll
            for (int i = 0; i < N; ++i) a[i] = t[i];
        }
    }
}

// calcSubFace for a face of exactly N points held in registers; the same
// operations in the same order as calcSubFace above.
// IF: record the plane crossings in out.ifPts (the cell cut needs them; the
// face-flux integrals only need the area).  out may live in LDS: it is then
// written in place, which keeps a per-lane copy of the record out of scratch.
// FULL: full holds face::centre and face::areaNormal of these points (6
// values), which a fully submerged face takes as they are instead of
// computing them again.  DOTS: dots holds dot(fPts[i], normal), so that a
// cut adds only the distance (the same values, computed once).
template <int N, bool IF = true, bool FULL = false, bool DOTS = FULL, class FC>
__device__ inline bool calcSubFaceN(const V3 (&fPts)[N], V3 normal, double distance, FC& out,
                                    const double* full = nullptr, const double* dots = nullptr) {
    auto fullCentre = [&]() {
        if constexpr (FULL) return v3(full[0], full[1], full[2]);
        else return faceCentre(fPts, N);
    };
    auto fullArea = [&]() {
        if constexpr (FULL) return v3(full[3], full[4], full[5]);
        else return faceAreaNormal(fPts, N);
    };
    double dist[N];
    int nSub = 0, first = -1;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double di;
        if constexpr (DOTS) di = dots[i] + distance;
        else di = dot(fPts[i], normal) + distance;
        if (fabs(di) < SVP_TSMALL) di += sgn(di) * SVP_TSMALL;
        dist[i] = di;
        if (di < 0.0) {
            nSub++;
            if (first == -1) first = i;
        }
    }
    out.nIf = 0;
    if (nSub == N || nSub == 0) {
        if (nSub == N) {
            out.status = -1;
            out.centre = fullCentre();
            out.area = fullArea();
        } else {
            out.status = 1;
            out.centre = VZERO;
            out.area = VZERO;
        }
        return true;
    }
    bool ok = true;
    // the walk starts at the first submerged point: rotate it to slot 0
    V3 rp[N];
    double rd[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        rp[i] = fPts[i];
        rd[i] = dist[i];
    }
    rotate_left(rp, first);
    rotate_left(rd, first);
    int nPts = 0;
    V3 p0 = VZERO, p1 = VZERO, p2 = VZERO, fCentre = VZERO;
    auto emit1 = [&](V3 q) {
        fCentre = vsel(nPts == 0, q, fCentre + q);
        p0 = vsel(nPts == 0, q, p0);
        p1 = vsel(nPts == 1, q, p1);
        p2 = vsel(nPts == 2, q, p2);
        nPts++;
    };
    static_assert(MAX_IF_POINTS == 4, "four crossing registers");
    V3 ip0 = VZERO, ip1 = VZERO, ip2 = VZERO, ip3 = VZERO;
    int nIf = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int nxt = (k + 1 == N) ? 0 : k + 1;
        const double dc = rd[k], dn = rd[nxt];
        const V3 pc = rp[k], pn = rp[nxt];
        if (dc < 0) emit1(pc);
        if ((dc * dn) < 0) {
            const double w = dc / (dc - dn);
            const V3 q = pc + w * (pn - pc);
            emit1(q);
            if constexpr (IF) {
                ip0 = vsel(nIf == 0, q, ip0);
                ip1 = vsel(nIf == 1, q, ip1);
                ip2 = vsel(nIf == 2, q, ip2);
                ip3 = vsel(nIf == 3, q, ip3);
            }
            if (nIf >= MAX_IF_POINTS) ok = false;
            nIf++;
        }
    }
    if constexpr (IF) {
        out.ifPts[0] = ip0;
        out.ifPts[1] = ip1;
        out.ifPts[2] = ip2;
        out.ifPts[3] = ip3;
    }
    out.nIf = nIf > MAX_IF_POINTS ? MAX_IF_POINTS : nIf;
    if (nPts < 3) {
        out.status = -1;
        out.centre = fullCentre();
        out.area = fullArea();
        return ok;
    }
    out.status = 0;
    if (nPts == 3) {
        out.centre = (1.0 / 3.0) * (p0 + p1 + p2);
        out.area = 0.5 * cross(p1 - p0, p2 - p0);
        return ok;
    }
    fCentre = fCentre / (double)nPts;
    V3 sumN = VZERO, sumAc = VZERO;
    double sumA = 0.0;
    V3 prev = VZERO, firstP = VZERO;
    int kk = 0;
    auto emit2 = [&](V3 q) {
        const bool on = kk > 0;
        firstP = vsel(on, firstP, q);
        if (on) {
            const V3 c = prev + q + fCentre;
            const V3 nn = cross(q - prev, fCentre - prev);
            const double ar = mag(nn);
            sumN = sumN + nn;
            sumA += ar;
            sumAc = sumAc + ar * c;
        }
        prev = q;
        kk++;
    };
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int nxt = (k + 1 == N) ? 0 : k + 1;
        const double dc = rd[k], dn = rd[nxt];
        const V3 pc = rp[k], pn = rp[nxt];
        if (dc < 0) emit2(pc);
        if ((dc * dn) < 0) {
            const double w = dc / (dc - dn);
            emit2(pc + w * (pn - pc));
        }
    }
    {
        const V3 c = prev + firstP + fCentre;
        const V3 nn = cross(firstP - prev, fCentre - prev);
        const double ar = mag(nn);
        sumN = sumN + nn;
        sumA += ar;
        sumAc = sumAc + ar * c;
    }
    if (sumA < SVP_ROOTVSMALL) {
        out.centre = fCentre;
        out.area = VZERO;
    } else {
        out.centre = (1.0 / 3.0) * sumAc / sumA;
        out.area = 0.5 * sumN;
    }
    return ok;
}

// Local description of one cell: its points and faces as point-index lists
// (cutCell::getLocalPointFieldAndFaceList, cutCell.C:140-236).  Without
// splitWarpedFace the faces are the mesh faces in their stored orientation.
template <int MF, int MP, int MI>
struct CellPolyT {
    static constexpr int maxFaces = MF, maxPts = MP, maxIdx = MI;
    int nPts, nFaces;
    V3 pts[MP];
    int fStart[MF + 1];
    unsigned char fIdx[MI];
};
using CellPolyNS = CellPolyT<NS_CELL_FACES, NS_CELL_POINTS, NS_FACE_IDX>;
using CellPolySP = CellPolyT<MAX_CELL_FACES, MAX_CELL_POINTS, SP_FACE_IDX>;
using CellPolyGen = CellPolyT<GEN_CELL_FACES, GEN_CELL_POINTS, GEN_FACE_IDX>;

// Result of cutting a cell (cutCell::calcSubCell)
struct CellCut {
    int status;
    double VOF;
    V3 ifCentre, ifArea;
};

template <class Poly>
__device__ inline void gatherFace(const Poly& P, int f, V3* buf, int& n) {
    n = P.fStart[f + 1] - P.fStart[f];
    for (int k = 0; k < n; ++k) buf[k] = P.pts[P.fIdx[P.fStart[f] + k]];
}

// cutCell::calcInterfaceCentreAndArea (cutCell.C:37-100) over face records
// in cell-face order; the orientation test is against the zero vector, which
// is what subCellCentre_ holds at that point of the reference.
__device__ __attribute__((always_inline)) inline void interfaceCentreAndArea(const FaceCut* fc, int nF, V3& ifC, V3& ifS) {
    V3 fCentre = VZERO;
    int nEdgePoints = 0;
    for (int f = 0; f < nF; ++f) {
        if (fc[f].status != 0) continue;
        for (int k = 0; k < fc[f].nIf; ++k) {
            fCentre = fCentre + fc[f].ifPts[k];
            nEdgePoints++;
        }
    }
    if (nEdgePoints > 0) fCentre = fCentre / (double)nEdgePoints;
    V3 sumN = VZERO, sumAc = VZERO;
    double sumA = 0.0;
    for (int f = 0; f < nF; ++f) {
        if (fc[f].status != 0) continue;
        for (int k = 0; k < fc[f].nIf - 1; ++k) {
            const V3 a = fc[f].ifPts[k], b = fc[f].ifPts[k + 1];
            const V3 c = a + b + fCentre;
            const V3 nn = cross(b - a, fCentre - a);
            const double ar = mag(nn);
            sumN = sumN + sgn(dot(nn, sumN)) * nn;
            sumA += ar;
            sumAc = sumAc + ar * c;
        }
    }
    if (sumA < SVP_ROOTVSMALL) {
        ifC = fCentre;
        ifS = VZERO;
    } else {
        ifC = (1.0 / 3.0) * sumAc / sumA;
        ifS = 0.5 * sumN;
    }
    if (dot(ifS, ifC - VZERO) < 0.0) ifS = (-1.0) * ifS;
}

// cutCell::calcSubCell (cutCell.C:343-542) combining face records fc[0..nF)
// produced in cell-face order.  Cc, Vc: cell centre and volume.
__device__ __attribute__((always_inline)) inline void combineSubCell(const FaceCut* fc, int nF, V3 Cc, double Vc, CellCut& out) {
    bool fullySubmerged = true, fullyEmpty = true;
    int nSubmergedFaces = 0;
    for (int f = 0; f < nF; ++f) {
        if (fc[f].status == 0) {
            fullySubmerged = false;
            fullyEmpty = false;
        } else if (fc[f].status == -1) {
            fullyEmpty = false;
            nSubmergedFaces++;
        } else {
            fullySubmerged = false;
        }
    }
    out.ifCentre = VZERO;
    out.ifArea = VZERO;
    if (!fullySubmerged && !fullyEmpty) {
        out.status = 0;
        interfaceCentreAndArea(fc, nF, out.ifCentre, out.ifArea);
        if (mag(out.ifArea) < SVP_TSMALL) {
            if (nSubmergedFaces == 0) {
                out.status = 1;
                out.VOF = 0.0;
            } else {
                out.status = -1;
                out.VOF = 1.0;
            }
            return;
        }
        // calcSubCellCentreAndVolume (cutCell.C:103-137): cut/submerged faces
        // in order, then the interface polygon
        V3 sum = VZERO;
        int cnt = 0;
        for (int f = 0; f < nF; ++f)
            if (fc[f].status != 1) { sum = sum + fc[f].centre; cnt++; }
        sum = sum + out.ifCentre;
        cnt++;
        const V3 cEst = sum / (double)cnt;
        double vol = 0.0;
        for (int f = 0; f < nF; ++f) {
            if (fc[f].status == 1) continue;
            const double pyr3Vol = fmax(fabs(dot(fc[f].area, fc[f].centre - cEst)), SVP_VSMALL);
            vol += pyr3Vol;
        }
        {
            const double pyr3Vol = fmax(fabs(dot(out.ifArea, out.ifCentre - cEst)), SVP_VSMALL);
            vol += pyr3Vol;
        }
        vol /= 3.0;
        out.VOF = vol / Vc;
    } else if (fullyEmpty) {
        out.status = 1;
        out.VOF = 0.0;
    } else {
        out.status = -1;
        out.VOF = 1.0;
    }
    (void)Cc;
}

// combineSubCell for a lane group that holds one face per lane (lane l: face
// l of the cell, records rec[0..nF) in LDS, written by their lanes).  The
// per-face terms that do not depend on the running sums are computed by the
// face's own lane: the first interface segment of a cut face (its centre sum,
// normal and area) and the pyramid volume against cEst; every lane then adds
// them in the reference's face order, so the sums and the result are those
// of combineSubCell.  seg/pyr: the group's scratch in LDS (nF entries each).
struct SegTerm {
    V3 c, nn;
    double ar;
};

__device__ __attribute__((always_inline)) inline void combineSubCellGroup(const FaceCut* fc, SegTerm* seg, double* pyr,
                                                                           int nF, int l, double Vc, CellCut& out) {
    bool fullySubmerged = true, fullyEmpty = true;
    int nSubmergedFaces = 0;
    for (int f = 0; f < nF; ++f) {
        if (fc[f].status == 0) {
            fullySubmerged = false;
            fullyEmpty = false;
        } else if (fc[f].status == -1) {
            fullyEmpty = false;
            nSubmergedFaces++;
        } else {
            fullySubmerged = false;
        }
    }
    out.ifCentre = VZERO;
    out.ifArea = VZERO;
    if (!fullySubmerged && !fullyEmpty) {
        out.status = 0;
        // calcInterfaceCentreAndArea (cutCell.C:37-100)
        V3 fCentre = VZERO;
        int nEdgePoints = 0;
        for (int f = 0; f < nF; ++f) {
            if (fc[f].status != 0) continue;
            for (int k = 0; k < fc[f].nIf; ++k) {
                fCentre = fCentre + fc[f].ifPts[k];
                nEdgePoints++;
            }
        }
        if (nEdgePoints > 0) fCentre = fCentre / (double)nEdgePoints;
        if (l < nF && fc[l].status == 0 && fc[l].nIf >= 2) {
            const V3 a = fc[l].ifPts[0], b = fc[l].ifPts[1];
            seg[l].c = a + b + fCentre;
            const V3 nn = cross(b - a, fCentre - a);
            seg[l].nn = nn;
            seg[l].ar = mag(nn);
        }
        __builtin_amdgcn_wave_barrier();
        V3 sumN = VZERO, sumAc = VZERO;
        double sumA = 0.0;
        for (int f = 0; f < nF; ++f) {
            if (fc[f].status != 0) continue;
            for (int k = 0; k < fc[f].nIf - 1; ++k) {
                V3 c, nn;
                double ar;
                if (k == 0) {
                    c = seg[f].c;
                    nn = seg[f].nn;
                    ar = seg[f].ar;
                } else {
                    const V3 a = fc[f].ifPts[k], b = fc[f].ifPts[k + 1];
                    c = a + b + fCentre;
                    nn = cross(b - a, fCentre - a);
                    ar = mag(nn);
                }
                sumN = sumN + sgn(dot(nn, sumN)) * nn;
                sumA += ar;
                sumAc = sumAc + ar * c;
            }
        }
        if (sumA < SVP_ROOTVSMALL) {
            out.ifCentre = fCentre;
            out.ifArea = VZERO;
        } else {
            out.ifCentre = (1.0 / 3.0) * sumAc / sumA;
            out.ifArea = 0.5 * sumN;
        }
        if (dot(out.ifArea, out.ifCentre - VZERO) < 0.0) out.ifArea = (-1.0) * out.ifArea;
        if (mag(out.ifArea) < SVP_TSMALL) {
            if (nSubmergedFaces == 0) {
                out.status = 1;
                out.VOF = 0.0;
            } else {
                out.status = -1;
                out.VOF = 1.0;
            }
            return;
        }
        // calcSubCellCentreAndVolume (cutCell.C:103-137)
        V3 sum = VZERO;
        int cnt = 0;
        for (int f = 0; f < nF; ++f)
            if (fc[f].status != 1) {
                sum = sum + fc[f].centre;
                cnt++;
            }
        sum = sum + out.ifCentre;
        cnt++;
        const V3 cEst = sum / (double)cnt;
        if (l < nF && fc[l].status != 1) pyr[l] = fmax(fabs(dot(fc[l].area, fc[l].centre - cEst)), SVP_VSMALL);
        __builtin_amdgcn_wave_barrier();
        double vol = 0.0;
        for (int f = 0; f < nF; ++f) {
            if (fc[f].status == 1) continue;
            vol += pyr[f];
        }
        {
            const double pyr3Vol = fmax(fabs(dot(out.ifArea, out.ifCentre - cEst)), SVP_VSMALL);
            vol += pyr3Vol;
        }
        vol /= 3.0;
        out.VOF = vol / Vc;
    } else if (fullyEmpty) {
        out.status = 1;
        out.VOF = 0.0;
    } else {
        out.status = -1;
        out.VOF = 1.0;
    }
}

// Cut the whole cell sequentially (one lane)
template <class Poly>
__device__ inline bool calcSubCell(const Poly& P, V3 Cc, double Vc, V3 normal, double distance,
                                   FaceCut* fc, CellCut& out) {
    bool ok = true;
    V3 buf[MAX_FACE_POINTS];
    for (int f = 0; f < P.nFaces; ++f) {
        int n;
        gatherFace(P, f, buf, n);
        ok &= calcSubFace(buf, n, normal, distance, fc[f]);
    }
    combineSubCell(fc, P.nFaces, Cc, Vc, out);
    return ok;
}

}  // namespace svp
