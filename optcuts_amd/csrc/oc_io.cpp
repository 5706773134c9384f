// OBJ mesh files and per-iteration trace lines in the formats OptCuts reads
// and writes (host code).
//
// Reading follows igl::readOBJ as main.cpp:1153 calls it (vendored
// libigl readOBJ.cpp): lines read one at a time; the first token names the
// record; "v" takes every number after the first character of the line, "vt"
// two or three numbers, "f" words of the forms i, i/t, i//n, i/t/n tried in
// that order, negative indices counting back from the current end of their
// list; comments, groups, smoothing and material lines are skipped, other
// records ignored. The lists must be rectangular (list_to_matrix); this
// reader further requires 3 vertex coordinates and triangles, which is what
// OptCuts' TriMesh consumes, and keeps the first two values of 3-value "vt"
// records. Numbers are read with strtod, which also accepts the inf, nan and
// hexadecimal forms a stream extraction rejects; lines have no length limit
// (igl::readOBJ splits lines longer than 2047 characters).
//
// Writing follows TriMesh::save (TriMesh.cpp:1758-1798): "v" lines with three
// std::scientific values (precision 6), "vt" lines with two, "f a/t b/t c/t"
// lines with 1-based indices, and TriMesh::saveAsMesh(path, F0, scaleUV)
// (TriMesh.cpp:1939-1969), which maps the cut mesh back to the input
// triangles F0. Trace lines follow Optimizer::writeEnergyValToFile and
// writeGradL2NormToFile (Optimizer.cpp:706-739): default stream formatting
// (precision 6, "%g").
#include "oc_io.hpp"

#include "optcuts_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg)
{
    g_err = msg;
    return OC_ERR_ARG;
}

struct ObjLists {
    std::vector<std::vector<double>> v, vt;
    std::vector<std::vector<long>> f, ft;
};

// The numbers a stream extraction of doubles reads from s: stops at the first
// token that does not start a number.
std::vector<double> numbers_from(const char* s)
{
    std::vector<double> out;
    const char* p = s;
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
        if (!*p) break;
        char* end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p) break;
        out.push_back(x);
        p = end;
    }
    return out;
}

int parse_obj(const char* path, ObjLists& L)
{
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(std::string("cannot open ") + path);
    std::string line;
    std::vector<char> buf(1 << 16);
    int lineNo = 0;
    int st = OC_OK;
    for (;;) {
        line.clear();
        bool any = false;
        while (std::fgets(buf.data(), (int)buf.size(), fp)) {
            any = true;
            line += buf.data();
            if (!line.empty() && line.back() == '\n') break;
        }
        if (!any) break;
        ++lineNo;
        char type[64];
        if (std::sscanf(line.c_str(), "%63s", type) != 1) continue;  // empty line
        const char* rest = line.c_str() + std::strlen(type);
        const std::string t(type);
        if (t == "v") {
            std::vector<double> x = numbers_from(line.c_str() + 1);
            if (x.size() < 3) {
                st = fail("vertex on line " + std::to_string(lineNo) + " has fewer than 3 coordinates");
                break;
            }
            L.v.push_back(std::move(x));
        } else if (t == "vn") {
            double x[3];
            if (std::sscanf(rest, "%lf %lf %lf", &x[0], &x[1], &x[2]) != 3) {
                st = fail("normal on line " + std::to_string(lineNo) + " needs 3 coordinates");
                break;
            }
        } else if (t == "vt") {
            double x[3];
            const int n = std::sscanf(rest, "%lf %lf %lf", &x[0], &x[1], &x[2]);
            if (n != 2 && n != 3) {
                st = fail("texture coordinate on line " + std::to_string(lineNo) + " needs 2 or 3 values");
                break;
            }
            L.vt.emplace_back(x, x + n);
        } else if (t == "f") {
            std::vector<long> fv, ftc;
            long nfn = 0;
            const long nv = (long)L.v.size(), ntc = (long)L.vt.size();
            auto shift = [](long i, long size) { return i < 0 ? i + size : i - 1; };
            const char* p = rest;
            char word[256];
            int off = 0;
            bool bad = false;
            while (std::sscanf(p, "%255s%n", word, &off) == 1) {
                p += off;
                long i, it, in;
                if (std::sscanf(word, "%ld/%ld/%ld", &i, &it, &in) == 3) {
                    fv.push_back(shift(i, nv));
                    ftc.push_back(shift(it, ntc));
                    ++nfn;
                } else if (std::sscanf(word, "%ld/%ld", &i, &it) == 2) {
                    fv.push_back(shift(i, nv));
                    ftc.push_back(shift(it, ntc));
                } else if (std::sscanf(word, "%ld//%ld", &i, &in) == 2) {
                    fv.push_back(shift(i, nv));
                    ++nfn;
                } else if (std::sscanf(word, "%ld", &i) == 1) {
                    fv.push_back(shift(i, nv));
                } else {
                    bad = true;
                    break;
                }
            }
            const size_t n = fv.size();
            const bool ok = n > 0 && (nfn == 0 || (size_t)nfn == n) && (ftc.empty() || ftc.size() == n);
            if (bad || !ok) {
                st = fail("face on line " + std::to_string(lineNo) + " has an invalid format");
                break;
            }
            L.f.push_back(std::move(fv));
            L.ft.push_back(std::move(ftc));
        }
        // comments, g, s, usemtl, mtllib and any other record: skipped
    }
    std::fclose(fp);
    return st;
}

struct ObjMesh {
    std::vector<double> V, UV;
    std::vector<int> F, FUV;
    int nV = 0, nUV = 0, nF = 0;
    bool hasFUV = false;
};

int load_obj(const char* path, ObjMesh& M)
{
    if (!path) return fail("null path");
    ObjLists L;
    if (int st = parse_obj(path, L)) return st;
    // rectangular lists (list_to_matrix)
    for (const auto& x : L.v)
        if (x.size() != L.v[0].size()) return fail("vertices have different numbers of coordinates");
    if (!L.v.empty() && L.v[0].size() != 3) return fail("vertices must have 3 coordinates");
    for (const auto& x : L.vt)
        if (x.size() != L.vt[0].size()) return fail("texture coordinates have different sizes");
    for (const auto& x : L.f)
        if (x.size() != L.f[0].size()) return fail("faces have different numbers of corners");
    if (!L.f.empty() && L.f[0].size() != 3) return fail("faces must be triangles");
    const bool withFt = !L.ft.empty() && !L.ft[0].empty();
    if (withFt)
        for (const auto& x : L.ft)
            if (x.size() != L.ft[0].size()) return fail("texture indices missing on some faces");
    M.nV = (int)L.v.size();
    M.nUV = (int)L.vt.size();
    M.nF = (int)L.f.size();
    M.hasFUV = withFt;
    M.V.resize(3 * (size_t)M.nV);
    for (int i = 0; i < M.nV; ++i)
        for (int j = 0; j < 3; ++j) M.V[3 * (size_t)i + j] = L.v[i][j];
    M.UV.resize(2 * (size_t)M.nUV);
    for (int i = 0; i < M.nUV; ++i)
        for (int j = 0; j < 2; ++j) M.UV[2 * (size_t)i + j] = L.vt[i][j];
    M.F.resize(3 * (size_t)M.nF);
    M.FUV.resize(withFt ? 3 * (size_t)M.nF : 0);
    for (int t = 0; t < M.nF; ++t)
        for (int j = 0; j < 3; ++j) {
            const long a = L.f[t][j];
            if (a < 0 || a >= M.nV) return fail("face " + std::to_string(t) + " has a vertex index out of range");
            M.F[3 * (size_t)t + j] = (int)a;
            if (withFt) {
                const long b = L.ft[t][j];
                if (b < 0 || b >= M.nUV)
                    return fail("face " + std::to_string(t) + " has a texture index out of range");
                M.FUV[3 * (size_t)t + j] = (int)b;
            }
        }
    return OC_OK;
}

int write_obj(const char* path, int nV, const double* V, int nF, const int* F, int nUV, const double* UV,
              const int* FUV)
{
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(std::string("cannot create ") + path);
    for (int i = 0; i < nV; ++i)
        std::fprintf(fp, "v %e %e %e\n", V[3 * (size_t)i], V[3 * (size_t)i + 1], V[3 * (size_t)i + 2]);
    for (int i = 0; i < nUV; ++i) std::fprintf(fp, "vt %e %e\n", UV[2 * (size_t)i], UV[2 * (size_t)i + 1]);
    const int* G = FUV ? FUV : F;
    for (int t = 0; t < nF; ++t) {
        const int* f = F + 3 * (size_t)t;
        const int* g = G + 3 * (size_t)t;
        std::fprintf(fp, "f %d/%d %d/%d %d/%d\n", f[0] + 1, g[0] + 1, f[1] + 1, g[1] + 1, f[2] + 1, g[2] + 1);
    }
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    return ok ? OC_OK : fail(std::string("write error on ") + path);
}

}  // namespace

namespace oc {
namespace io {

int trace_energy_line(FILE* fp, double Etotal, double Esd, double seam, double a0)
{
    return std::fprintf(fp, "%g %g %g %g\n", Etotal, Esd, seam, a0) < 0 ? -1 : 0;
}

int trace_grad_line(FILE* fp, double gsq, double gsqSD)
{
    return std::fprintf(fp, "%g %g\n", gsq, gsqSD) < 0 ? -1 : 0;
}

}  // namespace io
}  // namespace oc

extern "C" {

const char* oc_io_last_error(void) { return g_err.c_str(); }

int oc_obj_read_sizes(const char* path, int* nV, int* nUV, int* nF, int* hasFUV)
{
    ObjMesh M;
    if (int st = load_obj(path, M)) return st;
    if (nV) *nV = M.nV;
    if (nUV) *nUV = M.nUV;
    if (nF) *nF = M.nF;
    if (hasFUV) *hasFUV = M.hasFUV ? 1 : 0;
    return OC_OK;
}

int oc_obj_read(const char* path, double* V, double* UV, int* F, int* FUV)
{
    ObjMesh M;
    if (int st = load_obj(path, M)) return st;
    if (V) std::copy(M.V.begin(), M.V.end(), V);
    if (UV) std::copy(M.UV.begin(), M.UV.end(), UV);
    if (F) std::copy(M.F.begin(), M.F.end(), F);
    if (FUV && M.hasFUV) std::copy(M.FUV.begin(), M.FUV.end(), FUV);
    return OC_OK;
}

int oc_obj_write(const char* path, int nV, const double* V, int nF, const int* F, int nUV, const double* UV,
                 const int* FUV)
{
    if (!path || nV < 0 || nF < 0 || nUV < 0 || (nV && !V) || (nF && !F) || (nUV && !UV))
        return fail("oc_obj_write: bad arguments");
    for (long k = 0; k < 3L * nF; ++k) {
        if (F[k] < 0 || F[k] >= nV) return fail("oc_obj_write: vertex index out of range");
        if (FUV && (FUV[k] < 0 || FUV[k] >= nUV)) return fail("oc_obj_write: texture index out of range");
    }
    return write_obj(path, nV, V, nF, F, nUV, UV, FUV);
}

int oc_obj_write_mesh(const char* path, int nV, const double* Vrest, const double* UV, int nF, const int* F,
                      const int* F0, int scaleUV)
{
    if (!path || nV <= 0 || nF <= 0 || !Vrest || !UV || !F || !F0) return fail("oc_obj_write_mesh: bad arguments");
    int nV0 = 0;
    for (long k = 0; k < 3L * nF; ++k) {
        if (F[k] < 0 || F[k] >= nV) return fail("oc_obj_write_mesh: cut-mesh index out of range");
        if (F0[k] < 0) return fail("oc_obj_write_mesh: negative input-mesh index");
        nV0 = std::max(nV0, F0[k] + 1);
    }
    // input-mesh positions from the cut mesh's rest positions (last writer wins, as
    // the reference's row assignments do); rows no triangle names stay zero
    std::vector<double> V0(3 * (size_t)nV0, 0.0);
    for (long k = 0; k < 3L * nF; ++k)
        for (int j = 0; j < 3; ++j) V0[3 * (size_t)F0[k] + j] = Vrest[3 * (size_t)F[k] + j];
    std::vector<double> uv(UV, UV + 2 * (size_t)nV);
    if (scaleUV) {
        double umin = uv[0], umax = uv[0], vmin = uv[1], vmax = uv[1];
        for (int i = 1; i < nV; ++i) {
            umin = std::min(umin, uv[2 * (size_t)i]);
            umax = std::max(umax, uv[2 * (size_t)i]);
            vmin = std::min(vmin, uv[2 * (size_t)i + 1]);
            vmax = std::max(vmax, uv[2 * (size_t)i + 1]);
        }
        const double scale = std::max(umax - umin, vmax - vmin);
        for (int i = 0; i < nV; ++i) {
            uv[2 * (size_t)i] = (uv[2 * (size_t)i] - umin) / scale;
            uv[2 * (size_t)i + 1] = (uv[2 * (size_t)i + 1] - vmin) / scale;
        }
    }
    return write_obj(path, nV0, V0.data(), nF, F0, nV, uv.data(), F);
}

}  // extern "C"
