// OpenFOAM case files on the host: the polyMesh of constant/polyMesh
// (points, faces, owner, neighbour, boundary) and volFields (internalField
// and boundaryField entries), read and written in ASCII.  Each file may also
// be gzip-compressed as "<name>.gz", which OpenFOAM reads transparently and
// writes with writeCompression on.  The reader covers what fvMesh
// construction (polyMesh::polyMesh(IOobject)) and GeometricField's
// IOobject constructor take from these files for the implicit step: list
// sizes, the face lists (faceList or faceCompactList), the patch table with
// patch types and the cyclic pairing / transform, and field values given as
// `uniform` or `nonuniform List<...>`.
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "icsf.h"
#include "icsf_dict.h"

namespace icsf {
int reportError(int code, const std::string& msg);
}

using icsf::reportError;

namespace {

// file text: the file itself, else "<path>.gz"
bool readText(const std::string& path, std::string& out) {
    {
        std::ifstream f(path, std::ios::binary);
        if (f) {
            std::ostringstream ss;
            ss << f.rdbuf();
            out = ss.str();
            return true;
        }
    }
    gzFile g = gzopen((path + ".gz").c_str(), "rb");
    if (!g) return false;
    out.clear();
    char buf[1 << 16];
    int n;
    while ((n = gzread(g, buf, sizeof(buf))) > 0) out.append(buf, (size_t)n);
    const bool ok = n == 0;
    gzclose(g);
    return ok;
}

// Scanner over OpenFOAM ASCII: skips white space, // and /* */ comments.
struct Scan {
    const std::string& s;
    size_t i = 0;
    explicit Scan(const std::string& t) : s(t) {}
    void ws() {
        for (;;) {
            while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') ++i;
            } else if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                const size_t e = s.find("*/", i + 2);
                i = e == std::string::npos ? s.size() : e + 2;
            } else {
                return;
            }
        }
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { ++i; return true; }
        return false;
    }
    bool peek(char c) {
        ws();
        return i < s.size() && s[i] == c;
    }
    bool number(double& x) {
        ws();
        const char* b = s.c_str() + i;
        char* e = nullptr;
        x = std::strtod(b, &e);
        if (e == b) return false;
        i += (size_t)(e - b);
        return true;
    }
    bool integer(long& x) {
        ws();
        const char* b = s.c_str() + i;
        char* e = nullptr;
        x = std::strtol(b, &e, 10);
        if (e == b) return false;
        i += (size_t)(e - b);
        return true;
    }
    // skip the FoamFile { ... } header if present
    void header() {
        ws();
        if (s.compare(i, 8, "FoamFile") == 0) {
            int depth = 0;
            for (; i < s.size(); ++i) {
                if (s[i] == '{') ++depth;
                else if (s[i] == '}' && --depth == 0) { ++i; break; }
            }
        }
    }
};

// "N ( item ... )" or "N{value}" lists of scalars (n per item; vectors in
// parentheses when n == 3)
bool readList(Scan& sc, int n, std::vector<double>& out, std::string& err) {
    long N;
    if (!sc.integer(N) || N < 0) { err = "list size expected"; return false; }
    out.assign((size_t)N * n, 0.0);
    if (sc.eat('{')) {  // uniform list shorthand
        std::vector<double> v(n);
        const bool par = n > 1 && sc.eat('(');
        for (int k = 0; k < n; ++k)
            if (!sc.number(v[k])) { err = "list value expected"; return false; }
        if (par && !sc.eat(')')) { err = "')' expected"; return false; }
        if (!sc.eat('}')) { err = "'}' expected"; return false; }
        for (long j = 0; j < N; ++j)
            for (int k = 0; k < n; ++k) out[(size_t)j * n + k] = v[k];
        return true;
    }
    if (!sc.eat('(')) { err = "'(' expected"; return false; }
    for (long j = 0; j < N; ++j) {
        const bool par = n > 1;
        if (par && !sc.eat('(')) { err = "'(' expected at item " + std::to_string(j); return false; }
        for (int k = 0; k < n; ++k)
            if (!sc.number(out[(size_t)j * n + k])) { err = "number expected at item " + std::to_string(j); return false; }
        if (par && !sc.eat(')')) { err = "')' expected at item " + std::to_string(j); return false; }
    }
    if (!sc.eat(')')) { err = "')' expected at the end of the list"; return false; }
    return true;
}

bool readLabels(Scan& sc, std::vector<int>& out, std::string& err) {
    std::vector<double> v;
    if (!readList(sc, 1, v, err)) return false;
    out.resize(v.size());
    for (size_t j = 0; j < v.size(); ++j) out[j] = (int)v[j];
    return true;
}

// faceList "N ( n(v ...) ... )" or faceCompactList (offsets list, then
// vertices list)
bool readFaces(const std::string& text, std::vector<int>& off, std::vector<int>& verts, std::string& err) {
    Scan sc(text);
    sc.header();
    const bool compact = text.find("faceCompactList") != std::string::npos;
    if (compact) {
        if (!readLabels(sc, off, err) || !readLabels(sc, verts, err)) return false;
        return true;
    }
    long N;
    if (!sc.integer(N) || N < 0 || !sc.eat('(')) { err = "face list expected"; return false; }
    off.assign(1, 0);
    off.reserve((size_t)N + 1);
    verts.clear();
    verts.reserve((size_t)N * 4);
    for (long j = 0; j < N; ++j) {
        long n;
        if (!sc.integer(n) || n < 3 || !sc.eat('(')) { err = "face " + std::to_string(j) + " expected"; return false; }
        for (long k = 0; k < n; ++k) {
            long v;
            if (!sc.integer(v)) { err = "vertex label expected in face " + std::to_string(j); return false; }
            verts.push_back((int)v);
        }
        if (!sc.eat(')')) { err = "')' expected in face " + std::to_string(j); return false; }
        off.push_back((int)verts.size());
    }
    if (!sc.eat(')')) { err = "')' expected at the end of the faces"; return false; }
    return true;
}

struct PatchInfo {
    std::string name, type, nbrName;
    int start = 0, size = 0, kind = ICSF_PATCH, nbr = -1;
    double T[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::string transform;
    double axis[3] = {0, 0, 1};
    // phaseLagCyclic (phaseLagCyclicPolyPatch.C:79-90): IBPA, HBZoneName,
    // cylCoords (a required entry), the pair's rotationCentre
    bool phaseLag = false, cylCoords = false;
    double ibpa = 0, centre[3] = {0, 0, 0};
    std::string hbZone;
};

bool vec3(const std::vector<std::string>* v, double* out) {
    if (!v || v->empty()) return false;
    std::string t;
    for (const auto& s : *v) t += s + " ";
    for (char& c : t) if (c == '(' || c == ')') c = ' ';
    std::istringstream is(t);
    return bool(is >> out[0] >> out[1] >> out[2]);
}

// polyBoundaryMesh: "N ( name { type ...; nFaces ...; startFace ...; } ... )"
bool readBoundary(const std::string& text, std::vector<PatchInfo>& out, std::string& err) {
    Scan sc(text);
    sc.header();
    long N;
    if (!sc.integer(N) || N < 0 || !sc.eat('(')) { err = "boundary list expected"; return false; }
    for (long j = 0; j < N; ++j) {
        sc.ws();
        size_t b = sc.i;
        while (sc.i < text.size() && !std::isspace((unsigned char)text[sc.i]) && text[sc.i] != '{') ++sc.i;
        PatchInfo p;
        p.name = text.substr(b, sc.i - b);
        if (!sc.eat('{')) { err = "'{' expected after patch " + p.name; return false; }
        int depth = 1;
        b = sc.i;
        while (sc.i < text.size() && depth > 0) {
            if (text[sc.i] == '{') ++depth;
            else if (text[sc.i] == '}') --depth;
            ++sc.i;
        }
        icsf::Dict d;
        const std::string e = icsf::parseDict(text.substr(b, sc.i - 1 - b), d);
        if (!e.empty()) { err = "patch " + p.name + ": " + e; return false; }
        p.type = d.word("type", "patch");
        p.size = (int)d.scalar("nFaces", -1);
        p.start = (int)d.scalar("startFace", -1);
        if (p.size < 0 || p.start < 0) { err = "patch " + p.name + ": nFaces/startFace missing"; return false; }
        if (p.type == "wall") p.kind = ICSF_WALL;
        else if (p.type == "empty") p.kind = ICSF_EMPTY;
        else if (p.type == "cyclic" || p.type == "phaseLagCyclic") {
            p.kind = ICSF_CYCLIC;
            p.nbrName = d.word("neighbourPatch", "");
            p.transform = d.word("transform", d.word("transformType", "unknown"));
            vec3(d.find("rotationAxis"), p.axis);
            vec3(d.find("rotationCentre"), p.centre);
            if (p.type == "phaseLagCyclic") {
                p.phaseLag = true;
                p.ibpa = d.scalar("IBPA", 0.0);
                p.hbZone = d.word("HBZoneName", "");
                const std::string cc = d.word("cylCoords", "");
                if (cc == "true" || cc == "yes" || cc == "on" || cc == "1") p.cylCoords = true;
                else if (cc == "false" || cc == "no" || cc == "off" || cc == "0") p.cylCoords = false;
                else { err = "patch " + p.name + ": phaseLagCyclic needs cylCoords"; return false; }
            }
        } else if (p.type == "patch" || p.type == "symmetryPlane" || p.type == "symmetry") p.kind = ICSF_PATCH;
        else { err = "patch " + p.name + ": unsupported patch type " + p.type; return false; }
        out.push_back(p);
    }
    return true;
}

}  // namespace

struct icsf_foam_mesh {
    std::vector<double> points;
    std::vector<int> off, verts, owner, neighbour;
    std::vector<PatchInfo> patches;
    int nCells = 0;
};

struct icsf_foam_field {
    int nComp = 1;
    std::vector<double> internal;
    icsf::Dict dict;
};

namespace {

// forwardT of a rotational cyclic pair from the face normals: the rotation
// about the patch's rotationAxis that takes the neighbour patch's mean
// normal to minus this patch's mean normal (the coupled faces' normals are
// opposite once the far side is rotated into this side's frame);
// translational pairs and unknown transforms keep the identity
void cyclicTransforms(icsf_foam_mesh& m) {
    auto meanNormal = [&](const PatchInfo& p, double* n) {
        n[0] = n[1] = n[2] = 0;
        for (int f = p.start; f < p.start + p.size; ++f) {
            const int b = m.off[f], e = m.off[f + 1];
            const double* x0 = &m.points[3 * (size_t)m.verts[b]];
            for (int k = b + 1; k + 1 < e; ++k) {
                const double* x1 = &m.points[3 * (size_t)m.verts[k]];
                const double* x2 = &m.points[3 * (size_t)m.verts[k + 1]];
                const double a[3] = {x1[0] - x0[0], x1[1] - x0[1], x1[2] - x0[2]};
                const double c[3] = {x2[0] - x0[0], x2[1] - x0[1], x2[2] - x0[2]};
                n[0] += a[1] * c[2] - a[2] * c[1];
                n[1] += a[2] * c[0] - a[0] * c[2];
                n[2] += a[0] * c[1] - a[1] * c[0];
            }
        }
        const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int d = 0; d < 3; ++d) n[d] /= l;
    };
    for (auto& p : m.patches) {
        if (p.kind != ICSF_CYCLIC || p.transform != "rotational") continue;
        double a[3] = {p.axis[0], p.axis[1], p.axis[2]};
        const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (double& x : a) x /= la;
        double nP[3], nN[3];
        meanNormal(p, nP);
        meanNormal(m.patches[p.nbr], nN);
        // components perpendicular to the axis; angle from nN to -nP
        auto perp = [&](const double* v, double* o) {
            const double s = v[0] * a[0] + v[1] * a[1] + v[2] * a[2];
            for (int d = 0; d < 3; ++d) o[d] = v[d] - s * a[d];
        };
        double u[3], w[3];
        perp(nN, u);
        const double mnP[3] = {-nP[0], -nP[1], -nP[2]};
        perp(mnP, w);
        const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double sn = cr[0] * a[0] + cr[1] * a[1] + cr[2] * a[2];
        const double cs = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
        const double th = std::atan2(sn, cs);
        const double c = std::cos(th), s = std::sin(th), t = 1 - c;
        const double R[9] = {t * a[0] * a[0] + c,        t * a[0] * a[1] - s * a[2], t * a[0] * a[2] + s * a[1],
                             t * a[0] * a[1] + s * a[2], t * a[1] * a[1] + c,        t * a[1] * a[2] - s * a[0],
                             t * a[0] * a[2] - s * a[1], t * a[1] * a[2] + s * a[0], t * a[2] * a[2] + c};
        std::memcpy(p.T, R, sizeof(R));
    }
}

std::string fmtNum(double x) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", x);
    return b;
}

void foamHeader(std::ostream& o, const char* cls, const char* object, const std::string& note = "") {
    o << "/*--------------------------------*- C++ -*----------------------------------*\\\n"
         "  Written by the ICSFoam MI355X implicit-step library (OpenFOAM ASCII format)\n"
         "\\*---------------------------------------------------------------------------*/\n"
         "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       "
      << cls << ";\n";
    if (!note.empty()) o << "    note        \"" << note << "\";\n";
    o << "    object      " << object << ";\n}\n"
         "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n";
}

// field values of an entry: "uniform x", "uniform (x y z)", or
// "nonuniform List<...> N ( ... )"; n expected items, nComp per item
bool entryValues(const std::vector<std::string>& tok, int nComp, int n, std::vector<double>& out, std::string& err) {
    std::string t;
    for (const auto& s : tok) t += s + " ";
    Scan sc(t);
    sc.ws();
    if (t.compare(sc.i, 7, "uniform") == 0) {
        sc.i += 7;
        std::vector<double> v(nComp);
        const bool par = sc.eat('(');
        for (int k = 0; k < nComp; ++k)
            if (!sc.number(v[k])) { err = "uniform value expected"; return false; }
        if (par && !sc.eat(')')) { err = "')' expected"; return false; }
        out.resize((size_t)n * nComp);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nComp; ++k) out[(size_t)j * nComp + k] = v[k];
        return true;
    }
    if (t.compare(sc.i, 10, "nonuniform") == 0) {
        sc.i += 10;
        sc.ws();
        while (sc.i < t.size() && !std::isspace((unsigned char)t[sc.i])) ++sc.i;  // List<...>
        if (!readList(sc, nComp, out, err)) return false;
        if ((int)(out.size() / nComp) != n) {
            err = "list of " + std::to_string(out.size() / nComp) + " values where " + std::to_string(n) +
                  " are expected";
            return false;
        }
        return true;
    }
    err = "expected uniform or nonuniform values";
    return false;
}

}  // namespace

extern "C" {

int icsf_foam_mesh_read(const char* dir, icsf_foam_mesh** out) {
    if (!dir || !out) return reportError(ICSF_ERR_ARG, "icsf_foam_mesh_read: NULL argument");
    *out = nullptr;
    std::unique_ptr<icsf_foam_mesh> m(new icsf_foam_mesh);
    const std::string d(dir);
    std::string t, err;
    auto need = [&](const char* name) {
        if (!readText(d + "/" + name, t)) {
            err = std::string("cannot read ") + d + "/" + name + " (or .gz)";
            return false;
        }
        return true;
    };
    if (!need("points")) return reportError(ICSF_ERR_ARG, err);
    {
        Scan sc(t);
        sc.header();
        if (!readList(sc, 3, m->points, err)) return reportError(ICSF_ERR_ARG, "points: " + err);
    }
    if (!need("faces")) return reportError(ICSF_ERR_ARG, err);
    if (!readFaces(t, m->off, m->verts, err)) return reportError(ICSF_ERR_ARG, "faces: " + err);
    if (!need("owner")) return reportError(ICSF_ERR_ARG, err);
    {
        Scan sc(t);
        sc.header();
        if (!readLabels(sc, m->owner, err)) return reportError(ICSF_ERR_ARG, "owner: " + err);
    }
    if (!need("neighbour")) return reportError(ICSF_ERR_ARG, err);
    {
        Scan sc(t);
        sc.header();
        if (!readLabels(sc, m->neighbour, err)) return reportError(ICSF_ERR_ARG, "neighbour: " + err);
    }
    if (!need("boundary")) return reportError(ICSF_ERR_ARG, err);
    if (!readBoundary(t, m->patches, err)) return reportError(ICSF_ERR_ARG, "boundary: " + err);
    const int nF = (int)m->off.size() - 1, nI = (int)m->neighbour.size(), nP = (int)(m->points.size() / 3);
    if ((int)m->owner.size() != nF) return reportError(ICSF_ERR_ARG, "owner and faces differ in size");
    if (nI > nF) return reportError(ICSF_ERR_ARG, "more neighbours than faces");
    for (int v : m->verts)
        if (v < 0 || v >= nP) return reportError(ICSF_ERR_ARG, "face vertex label out of range");
    int nC = 0;
    for (int o : m->owner) nC = std::max(nC, o + 1);
    for (int o : m->neighbour) nC = std::max(nC, o + 1);
    m->nCells = nC;
    for (int f = 0; f < nF; ++f)
        if (m->owner[f] < 0) return reportError(ICSF_ERR_ARG, "negative owner label");
    int s = nI;
    for (auto& p : m->patches) {
        if (p.start != s) return reportError(ICSF_ERR_ARG, "patch " + p.name + " does not follow the previous one");
        s += p.size;
    }
    if (s != nF) return reportError(ICSF_ERR_ARG, "patches do not cover the boundary faces");
    for (auto& p : m->patches) {
        if (p.kind != ICSF_CYCLIC) continue;
        for (size_t q = 0; q < m->patches.size(); ++q)
            if (m->patches[q].name == p.nbrName) p.nbr = (int)q;
        if (p.nbr < 0) return reportError(ICSF_ERR_ARG, "cyclic patch " + p.name + ": neighbourPatch not found");
        if (m->patches[p.nbr].size != p.size)
            return reportError(ICSF_ERR_ARG, "cyclic patch " + p.name + ": sizes of the pair differ");
    }
    cyclicTransforms(*m);
    *out = m.release();
    return ICSF_OK;
}

int icsf_foam_mesh_sizes(const icsf_foam_mesh* m, int* sizes) {
    if (!m || !sizes) return reportError(ICSF_ERR_ARG, "icsf_foam_mesh_sizes: NULL argument");
    sizes[0] = (int)(m->points.size() / 3);
    sizes[1] = (int)m->off.size() - 1;
    sizes[2] = (int)m->neighbour.size();
    sizes[3] = m->nCells;
    sizes[4] = (int)m->patches.size();
    sizes[5] = (int)m->verts.size();
    return ICSF_OK;
}

int icsf_foam_mesh_get(const icsf_foam_mesh* m, double* points, int* faceOffsets, int* faceVerts, int* owner,
                       int* neighbour, int* patchStart, int* patchSize, int* patchKind, int* patchNeighbour,
                       double* patchTransform) {
    if (!m) return reportError(ICSF_ERR_ARG, "NULL mesh");
    if (points) std::memcpy(points, m->points.data(), m->points.size() * sizeof(double));
    if (faceOffsets) std::memcpy(faceOffsets, m->off.data(), m->off.size() * sizeof(int));
    if (faceVerts) std::memcpy(faceVerts, m->verts.data(), m->verts.size() * sizeof(int));
    if (owner) std::memcpy(owner, m->owner.data(), m->owner.size() * sizeof(int));
    if (neighbour) std::memcpy(neighbour, m->neighbour.data(), m->neighbour.size() * sizeof(int));
    for (size_t p = 0; p < m->patches.size(); ++p) {
        const PatchInfo& q = m->patches[p];
        if (patchStart) patchStart[p] = q.start;
        if (patchSize) patchSize[p] = q.size;
        if (patchKind) patchKind[p] = q.kind;
        if (patchNeighbour) patchNeighbour[p] = q.nbr;
        if (patchTransform)
            for (int k = 0; k < 9; ++k) patchTransform[9 * p + k] = q.kind == ICSF_CYCLIC ? q.T[k] : 0.0;
    }
    return ICSF_OK;
}

int icsf_foam_mesh_patch_name(const icsf_foam_mesh* m, int patch, char* name, int len) {
    if (!m || !name || len < 1 || patch < 0 || patch >= (int)m->patches.size())
        return reportError(ICSF_ERR_ARG, "icsf_foam_mesh_patch_name: bad argument");
    std::snprintf(name, (size_t)len, "%s", m->patches[patch].name.c_str());
    return ICSF_OK;
}

int icsf_foam_mesh_patch_phase_lag(const icsf_foam_mesh* m, int patch, int* phaseLag, double* ibpa, int* cylCoords,
                                   double* rotationAxis, double* rotationCentre, char* hbZoneName, int len) {
    if (!m || !phaseLag || patch < 0 || patch >= (int)m->patches.size())
        return reportError(ICSF_ERR_ARG, "icsf_foam_mesh_patch_phase_lag: bad argument");
    const PatchInfo& p = m->patches[patch];
    *phaseLag = p.phaseLag ? 1 : 0;
    if (ibpa) *ibpa = p.ibpa;
    if (cylCoords) *cylCoords = p.cylCoords ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        if (rotationAxis) rotationAxis[d] = p.axis[d];
        if (rotationCentre) rotationCentre[d] = p.centre[d];
    }
    if (hbZoneName && len > 0) std::snprintf(hbZoneName, (size_t)len, "%s", p.hbZone.c_str());
    return ICSF_OK;
}

int icsf_foam_mesh_free(icsf_foam_mesh* m) {
    delete m;
    return ICSF_OK;
}

int icsf_foam_mesh_write(const char* dir, int nPoints, const double* points, int nFaces, const int* faceOffsets,
                         const int* faceVerts, int nInternalFaces, const int* owner, const int* neighbour,
                         int nPatches, const char* const* patchNames, const int* patchStart, const int* patchSize,
                         const int* patchKind, const int* patchNeighbour, const double* patchTransform) {
    if (!dir || !points || !faceOffsets || !faceVerts || !owner || (nInternalFaces > 0 && !neighbour) ||
        (nPatches > 0 && (!patchNames || !patchStart || !patchSize || !patchKind)))
        return reportError(ICSF_ERR_ARG, "icsf_foam_mesh_write: NULL argument");
    const std::string d(dir);
    int nC = 0;
    for (int f = 0; f < nFaces; ++f) nC = std::max(nC, owner[f] + 1);
    for (int f = 0; f < nInternalFaces; ++f) nC = std::max(nC, neighbour[f] + 1);
    const std::string note = "nPoints:" + std::to_string(nPoints) + "  nCells:" + std::to_string(nC) +
                             "  nFaces:" + std::to_string(nFaces) + "  nInternalFaces:" + std::to_string(nInternalFaces);
    auto open = [&](const char* name, std::ofstream& o) {
        o.open(d + "/" + name);
        return bool(o);
    };
    std::ofstream o;
    if (!open("points", o)) return reportError(ICSF_ERR_ARG, "cannot write " + d + "/points");
    foamHeader(o, "vectorField", "points");
    o << nPoints << "\n(\n";
    for (int i = 0; i < nPoints; ++i)
        o << "(" << fmtNum(points[3 * i]) << " " << fmtNum(points[3 * i + 1]) << " " << fmtNum(points[3 * i + 2]) << ")\n";
    o << ")\n";
    o.close();
    if (!open("faces", o)) return reportError(ICSF_ERR_ARG, "cannot write " + d + "/faces");
    foamHeader(o, "faceList", "faces");
    o << nFaces << "\n(\n";
    for (int f = 0; f < nFaces; ++f) {
        o << (faceOffsets[f + 1] - faceOffsets[f]) << "(";
        for (int k = faceOffsets[f]; k < faceOffsets[f + 1]; ++k) o << (k > faceOffsets[f] ? " " : "") << faceVerts[k];
        o << ")\n";
    }
    o << ")\n";
    o.close();
    for (int which = 0; which < 2; ++which) {
        const char* name = which ? "neighbour" : "owner";
        if (!open(name, o)) return reportError(ICSF_ERR_ARG, std::string("cannot write ") + d + "/" + name);
        foamHeader(o, "labelList", name, note);
        const int n = which ? nInternalFaces : nFaces;
        const int* a = which ? neighbour : owner;
        o << n << "\n(\n";
        for (int i = 0; i < n; ++i) o << a[i] << "\n";
        o << ")\n";
        o.close();
    }
    if (!open("boundary", o)) return reportError(ICSF_ERR_ARG, "cannot write " + d + "/boundary");
    foamHeader(o, "polyBoundaryMesh", "boundary");
    o << nPatches << "\n(\n";
    for (int p = 0; p < nPatches; ++p) {
        static const char* types[] = {"patch", "wall", "empty", "cyclic"};
        const int k = patchKind[p];
        if (k < 0 || k > 3) return reportError(ICSF_ERR_ARG, "unknown patch kind");
        o << "    " << patchNames[p] << "\n    {\n        type            " << types[k] << ";\n";
        o << "        nFaces          " << patchSize[p] << ";\n        startFace       " << patchStart[p] << ";\n";
        if (k == ICSF_CYCLIC) {
            if (!patchNeighbour || patchNeighbour[p] < 0 || patchNeighbour[p] >= nPatches)
                return reportError(ICSF_ERR_ARG, "cyclic patch without a neighbour");
            o << "        neighbourPatch  " << patchNames[patchNeighbour[p]] << ";\n";
            const double* T = patchTransform ? patchTransform + 9 * p : nullptr;
            bool ident = true;
            if (T)
                for (int q = 0; q < 9; ++q) ident = ident && std::fabs(T[q] - (q % 4 == 0 ? 1.0 : 0.0)) < 1e-14;
            if (ident) {
                o << "        transform       translational;\n";
            } else {
                // rotation axis: the eigenvector of T for eigenvalue 1
                const double ax[3] = {T[7] - T[5], T[2] - T[6], T[3] - T[1]};
                o << "        transform       rotational;\n        rotationAxis    (" << fmtNum(ax[0]) << " "
                  << fmtNum(ax[1]) << " " << fmtNum(ax[2]) << ");\n        rotationCentre  (0 0 0);\n";
            }
        }
        o << "    }\n";
    }
    o << ")\n";
    o.close();
    return ICSF_OK;
}

int icsf_foam_field_read(const char* path, int nCells, icsf_foam_field** out) {
    if (!path || !out || nCells < 0) return reportError(ICSF_ERR_ARG, "icsf_foam_field_read: bad argument");
    *out = nullptr;
    std::string t, err;
    if (!readText(path, t)) return reportError(ICSF_ERR_ARG, std::string("cannot read ") + path + " (or .gz)");
    std::unique_ptr<icsf_foam_field> f(new icsf_foam_field);
    const std::string e = icsf::parseDict(t, f->dict);
    if (!e.empty()) return reportError(ICSF_ERR_ARG, std::string(path) + ": " + e);
    const std::string cls = f->dict.word("FoamFile/class", "");
    if (cls == "volScalarField") f->nComp = 1;
    else if (cls == "volVectorField") f->nComp = 3;
    else return reportError(ICSF_ERR_ARG, std::string(path) + ": class " + cls + " is not a volScalarField or volVectorField");
    const auto* in = f->dict.find("internalField");
    if (!in) return reportError(ICSF_ERR_ARG, std::string(path) + ": no internalField");
    if (!entryValues(*in, f->nComp, nCells, f->internal, err))
        return reportError(ICSF_ERR_ARG, std::string(path) + ": internalField: " + err);
    if (!f->dict.sub("boundaryField")) return reportError(ICSF_ERR_ARG, std::string(path) + ": no boundaryField");
    *out = f.release();
    return ICSF_OK;
}

int icsf_foam_field_internal(const icsf_foam_field* f, int* nComp, double* values) {
    if (!f) return reportError(ICSF_ERR_ARG, "NULL field");
    if (nComp) *nComp = f->nComp;
    if (values) std::memcpy(values, f->internal.data(), f->internal.size() * sizeof(double));
    return ICSF_OK;
}

int icsf_foam_field_patch_type(const icsf_foam_field* f, const char* patch, char* type, int len) {
    if (!f || !patch || !type || len < 1) return reportError(ICSF_ERR_ARG, "icsf_foam_field_patch_type: bad argument");
    const icsf::Dict* b = f->dict.sub("boundaryField");
    const icsf::Dict* p = b ? b->sub(patch) : nullptr;
    if (!p) return reportError(ICSF_ERR_ARG, std::string("boundaryField has no entry for patch ") + patch);
    std::snprintf(type, (size_t)len, "%s", p->word("type", "").c_str());
    return ICSF_OK;
}

int icsf_foam_field_patch_entry(const icsf_foam_field* f, const char* patch, const char* key, int nFaces,
                                double* values) {
    if (!f || !patch || !key || nFaces < 0) return reportError(ICSF_ERR_ARG, "icsf_foam_field_patch_entry: bad argument");
    const icsf::Dict* b = f->dict.sub("boundaryField");
    const icsf::Dict* p = b ? b->sub(patch) : nullptr;
    if (!p) return reportError(ICSF_ERR_ARG, std::string("boundaryField has no entry for patch ") + patch);
    const auto* v = p->find(key);
    if (!v) return reportError(ICSF_ERR_ARG, std::string("patch ") + patch + " has no entry " + key);
    std::vector<std::string> tok = *v;
    // $internalField: the uniform internal value (OpenFOAM macro expansion)
    if (tok.size() == 1 && tok[0] == "$internalField") {
        const auto* in = f->dict.find("internalField");
        tok = *in;
    }
    std::vector<double> out;
    std::string err;
    if (!entryValues(tok, f->nComp, nFaces, out, err))
        return reportError(ICSF_ERR_ARG, std::string("patch ") + patch + ", entry " + key + ": " + err);
    if (values) std::memcpy(values, out.data(), out.size() * sizeof(double));
    return ICSF_OK;
}

int icsf_foam_field_free(icsf_foam_field* f) {
    delete f;
    return ICSF_OK;
}

int icsf_foam_field_write(const char* path, const char* name, const char* dimensions, int nComp, int nCells,
                          const double* internal, int nPatches, const char* const* patchNames,
                          const char* const* patchTypes, const int* patchSizes, const double* const* patchValues,
                          const char* const* patchEntries) {
    if (!path || !name || !internal || (nComp != 1 && nComp != 3) || nCells < 0 ||
        (nPatches > 0 && (!patchNames || !patchTypes || !patchSizes)))
        return reportError(ICSF_ERR_ARG, "icsf_foam_field_write: bad argument");
    std::ofstream o(path);
    if (!o) return reportError(ICSF_ERR_ARG, std::string("cannot write ") + path);
    auto item = [&](const double* v) {
        if (nComp == 1) return fmtNum(v[0]);
        return "(" + fmtNum(v[0]) + " " + fmtNum(v[1]) + " " + fmtNum(v[2]) + ")";
    };
    const char* tn = nComp == 1 ? "scalar" : "vector";
    foamHeader(o, nComp == 1 ? "volScalarField" : "volVectorField", name);
    o << "dimensions      " << (dimensions ? dimensions : "[0 0 0 0 0 0 0]") << ";\n\n";
    o << "internalField   nonuniform List<" << tn << ">\n" << nCells << "\n(\n";
    for (int i = 0; i < nCells; ++i) o << item(internal + (size_t)nComp * i) << "\n";
    o << ")\n;\n\nboundaryField\n{\n";
    for (int p = 0; p < nPatches; ++p) {
        o << "    " << patchNames[p] << "\n    {\n        type            " << patchTypes[p] << ";\n";
        if (patchEntries && patchEntries[p]) o << patchEntries[p];
        const double* v = patchValues ? patchValues[p] : nullptr;
        if (v) {
            o << "        value           nonuniform List<" << tn << "> " << patchSizes[p] << "(";
            for (int i = 0; i < patchSizes[p]; ++i) o << (i ? " " : "") << item(v + (size_t)nComp * i);
            o << ");\n";
        }
        o << "    }\n";
    }
    o << "}\n";
    return o ? ICSF_OK : reportError(ICSF_ERR_ARG, std::string("write failed: ") + path);
}

}  // extern "C"
