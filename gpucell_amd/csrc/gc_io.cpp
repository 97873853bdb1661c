// gc_io.cpp -- the reference's case decks and Tecplot frames, host only.
//
// Readers/writers for the four fixed-width (20-character column) input files
// of Particle::LoadParInfo (src/Particle.cu:45-427) and the six Tecplot
// outputs of Particle::OutputPar (src/Particle.cu:681-1300), with the same
// file names, line structure, field widths and number formats, so GPUCell
// case directories and post-processing scripts work unchanged.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gpucell/gpucell.h"

namespace gc {
int set_error(int code, const std::string& msg);
}

namespace {

// ---- fixed-width reading ------------------------------------------------
struct Deck {
    std::ifstream in;
    std::string path;
    int line = 0;
    bool ok = true;

    bool next(std::string& s)
    {
        if (!std::getline(in, s)) { ok = false; return false; }
        ++line;
        if (!s.empty() && s.back() == '\r') s.pop_back();
        return true;
    }
    // column c (0-based) of 20 characters
    static std::string col(const std::string& s, int c)
    {
        size_t b = (size_t)c * 20;
        return b < s.size() ? s.substr(b, 20) : std::string();
    }
    bool num(const std::string& s, int c, double& v)
    {
        std::string f = col(s, c);
        char* end = nullptr;
        const char* p = f.c_str();
        v = std::strtod(p, &end);
        return end != p;
    }
};

int open_deck(Deck& d, const std::string& dir, const char* name)
{
    d.path = dir + "/" + name;
    d.in.open(d.path);
    if (!d.in) return gc::set_error(GC_ERR_INVALID, "cannot open " + d.path);
    return GC_OK;
}

int bad(const Deck& d, const char* what)
{
    return gc::set_error(GC_ERR_INVALID, d.path + ":" + std::to_string(d.line) + ": " + what);
}

int read_values(Deck& d, int count, std::vector<double>& out, int ncol)
{
    std::string s;
    for (int i = 0; i < count; ++i) {
        if (!d.next(s)) return bad(d, "unexpected end of file");
        for (int c = 0; c < ncol; ++c) {
            double v;
            if (!d.num(s, c, v)) return bad(d, "missing number");
            out.push_back(v);
        }
    }
    return GC_OK;
}

template <class T> T* dup(const std::vector<T>& v)
{
    T* p = (T*)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

// ---- fixed-width writing ------------------------------------------------
// Numbers are written left-aligned in 20 columns with at most 16
// characters, so every field ends in blanks (the reference copies exactly
// 20 characters into a 20-byte buffer before atof, Particle.cu:50-51).
void put(std::ostream& o, double v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%-20.9g", v);
    o << b;
}
void put_int(std::ostream& o, long v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%-20ld", v);
    o << b;
}

int mkdirs(const std::string& path)
{
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            return gc::set_error(GC_ERR_INVALID, "cannot create " + cur);
    }
    return GC_OK;
}

}  // namespace

// -------------------------------------------------------------------------
extern "C" int gc_ReadCase(const char* dir_c, gc_case_desc* out)
{
    if (!dir_c || !out) return gc::set_error(GC_ERR_INVALID, "ReadCase: null argument");
    std::string dir(dir_c);
    std::memset(out, 0, sizeof(*out));
    int st;
    std::string s;
    Structconst& c = out->consts;
    // SDPDParCoor.txt (Particle::LoadFBPar, Particle.cu:45-122)
    std::vector<double> xyz;
    {
        Deck d;
        if ((st = open_deck(d, dir, "SDPDParCoor.txt"))) return st;
        double n1, n2, nf;
        if (!d.next(s) || !d.next(s) || !d.num(s, 0, n1) || !d.num(s, 1, n2) || !d.num(s, 2, nf))
            return bad(d, "expected nBound1 | nBound2 | nFluid");
        out->nBound1Par = (int)n1; out->nBound2Par = (int)n2; out->nFluidPar = (int)nf;
        const int cnt[3] = {out->nBound1Par, out->nBound2Par, out->nFluidPar};
        for (int b = 0; b < 3; ++b) {
            if (!d.next(s)) return bad(d, "missing block header");
            if ((st = read_values(d, cnt[b], xyz, 3))) return st;
        }
    }
    // SDPDPhyPara.txt (Particle::LoadSDPDPara, Particle.cu:124-172)
    {
        Deck d;
        if ((st = open_deck(d, dir, "SDPDPhyPara.txt"))) return st;
        double v[6];
        if (!d.next(s) || !d.next(s) || !d.num(s, 0, v[0])) return bad(d, "SoundSpeed1");
        if (!d.next(s) || !d.num(s, 0, v[1])) return bad(d, "SoundSpeed2");
        for (int k = 2; k < 6; ++k)
            if (!d.next(s) || !d.next(s) || !d.num(s, 0, v[k])) return bad(d, "parameter");
        c.SoundSpeed1 = (float)v[0]; c.SoundSpeed2 = (float)v[1];
        c.ShearVis = (float)v[2]; c.BulkVis = (float)v[3]; c.DenInit = (float)v[4];
        c.ConstExtForce = (float)v[5];  // times Mass below
    }
    // compile-time constants of main.cuh:6-16 and AllocCpuMemory (Particle.cu:436-437)
    c.TimeStep = 1.0e-3f; c.Mass = 1.0f; c.Temp = 1.0f; c.CutRadius = 1.0f;
    c.VVsigma = 0.65f; c.iSeed = 1;
    c.ConstExtForce *= c.Mass;
    const float bbox[6] = {0, 20, -10, 10, -10, 10};
    std::memcpy(out->bbox, bbox, sizeof bbox);
    gc_DefaultExtConsts(&out->ext);
    // CELLParStruct.txt (Particle::LoadCellStruct, Particle.cu:174-303)
    std::vector<double> cxyz;
    std::vector<tint3> tri;
    std::vector<int6> edge;
    std::vector<int> celltab, type;
    std::vector<tint2> betri;
    {
        Deck d;
        if ((st = open_deck(d, dir, "CELLParStruct.txt"))) return st;
        double nc;
        if (!d.next(s) || !d.next(s) || !d.num(s, 0, nc)) return bad(d, "nCell");
        int vb = 0, tb = 0;
        for (int i = 0; i < (int)nc; ++i) {
            double ty, np, nt, ne;
            if (!d.next(s) || !d.next(s) || !d.num(s, 0, ty) || !d.num(s, 1, np) || !d.num(s, 2, nt))
                return bad(d, "Type | nPar | nTri | nEdgeInfo");
            (void)ne;
            type.push_back((int)ty);
            if (!d.next(s)) return bad(d, "coordinate header");
            if ((st = read_values(d, (int)np, cxyz, 3))) return st;
            for (int k = 0; k < (int)np; ++k) celltab.push_back(i);
            if (!d.next(s)) return bad(d, "triangle header");
            std::vector<double> t;
            if ((st = read_values(d, (int)nt, t, 3))) return st;
            for (int k = 0; k < (int)nt; ++k)
                tri.push_back(tint3{(int)t[3 * k] + vb, (int)t[3 * k + 1] + vb, (int)t[3 * k + 2] + vb});
            int nedge = (int)nt * 3 / 2;
            if (!d.next(s)) return bad(d, "edge header");
            std::vector<double> e;
            if ((st = read_values(d, nedge, e, 6))) return st;
            for (int k = 0; k < nedge; ++k)
                edge.push_back(int6{(int)e[6 * k] + vb, (int)e[6 * k + 1] + vb, (int)e[6 * k + 2] + vb,
                                    (int)e[6 * k + 3] + vb, (int)e[6 * k + 4] + tb, (int)e[6 * k + 5] + tb});
            betri.push_back(tint2{tb, tb + (int)nt - 1});
            tb += (int)nt;
            vb += (int)np;
        }
        out->nCell = (int)nc;
        out->nAllTri = (int)tri.size();
        out->nAllEdge = (int)edge.size();
        out->nCellPar = (int)celltab.size();
    }
    // CELLPhyPara.txt (Particle::LoadCellPara, Particle.cu:305-427)
    std::vector<float> mu(out->nAllEdge), kc(out->nAllEdge), kd(out->nAllTri), ka(out->nAllTri),
        kv(out->nAllTri);
    if (out->nCell > 0) {
        Deck d;
        if ((st = open_deck(d, dir, "CELLPhyPara.txt"))) return st;
        double nt;
        if (!d.next(s) || !d.next(s) || !d.num(s, 0, nt)) return bad(d, "nType");
        std::vector<double> tab[5];
        for (int k = 0; k < 5; ++k) {
            if (!d.next(s)) return bad(d, "modulus header");
            if ((st = read_values(d, (int)nt, tab[k], 1))) return st;
        }
        double mc[4];
        for (int k = 0; k < 4; ++k)
            if (!d.next(s) || !d.next(s) || !d.num(s, 0, mc[k])) return bad(d, "Morse parameter");
        c.MaxCut = (float)mc[0]; c.zeroLen = (float)mc[1]; c.scaFac = (float)mc[2];
        c.surEnergy = (float)mc[3];
        // per-type -> per-edge / per-triangle broadcast (Particle.cu:405-425)
        int e = 0, t = 0;
        for (int i = 0; i < out->nCell; ++i) {
            int ty = type[i] - 1;
            if (ty < 0 || ty >= (int)nt) return bad(d, "cell type out of range");
            int ntri = betri[i].y - betri[i].x + 1;
            for (int k = 0; k < ntri * 3 / 2; ++k, ++e) { mu[e] = (float)tab[0][ty]; kc[e] = (float)tab[1][ty]; }
            for (int k = 0; k < ntri; ++k, ++t) {
                ka[t] = (float)tab[2][ty]; kd[t] = (float)tab[3][ty]; kv[t] = (float)tab[4][ty];
            }
        }
    }
    // CoorMerge (Particle.cu:462-469): SDPD particles then membrane vertices
    std::vector<tfloat3> coor;
    for (size_t i = 0; i < xyz.size() / 3; ++i)
        coor.push_back(tfloat3{(float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2]});
    for (size_t i = 0; i < cxyz.size() / 3; ++i)
        coor.push_back(tfloat3{(float)cxyz[3 * i], (float)cxyz[3 * i + 1], (float)cxyz[3 * i + 2]});
    c.nBound2Start = (unsigned)out->nBound1Par;
    c.nFluidStart = c.nBound2Start + (unsigned)out->nBound2Par;
    c.nCellStart = c.nFluidStart + (unsigned)out->nFluidPar;
    c.nCellPar = (unsigned)out->nCellPar;
    c.nTotalPar = c.nCellStart + c.nCellPar;
    out->coor = dup(coor);
    out->vel = nullptr;
    out->tri = dup(tri);
    out->edge = dup(edge);
    out->celltab = dup(celltab);
    out->beginendtri = dup(betri);
    out->shearmod_para = dup(mu);
    out->bendingmod_para = dup(kc);
    out->hlarea_para = dup(kd);
    out->hgarea_para = dup(ka);
    out->hvol_para = dup(kv);
    return GC_OK;
}

extern "C" void gc_FreeCase(gc_case_desc* d)
{
    if (!d) return;
    std::free((void*)d->coor); std::free((void*)d->vel); std::free((void*)d->tri);
    std::free((void*)d->edge); std::free((void*)d->celltab); std::free((void*)d->beginendtri);
    std::free((void*)d->shearmod_para); std::free((void*)d->bendingmod_para);
    std::free((void*)d->hlarea_para); std::free((void*)d->hgarea_para); std::free((void*)d->hvol_para);
    std::memset(d, 0, sizeof(*d));
}

// Writes the four decks.  Moduli are written per cell type; cells whose
// per-edge/per-triangle values are uniform and equal share one type.
extern "C" int gc_WriteCase(const char* dir_c, const gc_case_desc* d)
{
    if (!dir_c || !d || !d->coor) return gc::set_error(GC_ERR_INVALID, "WriteCase: null argument");
    std::string dir(dir_c);
    int st = mkdirs(dir);
    if (st) return st;
    const Structconst& c = d->consts;
    const int nsdpd = d->nBound1Par + d->nBound2Par + d->nFluidPar;
    {
        std::ofstream o(dir + "/SDPDParCoor.txt");
        if (!o) return gc::set_error(GC_ERR_INVALID, "cannot write SDPDParCoor.txt");
        o << "nBound1 nBound2 nFluid\n";
        put_int(o, d->nBound1Par); put_int(o, d->nBound2Par); put_int(o, d->nFluidPar); o << "\n";
        const char* hdr[3] = {"Bound1 x y z", "Bound2 x y z", "Fluid x y z"};
        int b0 = 0;
        const int cnt[3] = {d->nBound1Par, d->nBound2Par, d->nFluidPar};
        for (int b = 0; b < 3; ++b) {
            o << hdr[b] << "\n";
            for (int i = b0; i < b0 + cnt[b]; ++i) {
                put(o, d->coor[i].x); put(o, d->coor[i].y); put(o, d->coor[i].z); o << "\n";
            }
            b0 += cnt[b];
        }
    }
    {
        std::ofstream o(dir + "/SDPDPhyPara.txt");
        if (!o) return gc::set_error(GC_ERR_INVALID, "cannot write SDPDPhyPara.txt");
        o << "SoundSpeed1 / SoundSpeed2\n";
        put(o, c.SoundSpeed1); o << "\n"; put(o, c.SoundSpeed2); o << "\n";
        o << "ShearVis\n"; put(o, c.ShearVis); o << "\n";
        o << "BulkVis\n"; put(o, c.BulkVis); o << "\n";
        o << "DenInit\n"; put(o, c.DenInit); o << "\n";
        o << "ConstExtForce (per unit mass)\n"; put(o, c.ConstExtForce / c.Mass); o << "\n";
    }
    // cell types: one per distinct (mu, kc, ka, kd, kv) of a cell's first element
    std::vector<std::vector<float>> types;
    std::vector<int> celltype(d->nCell);
    {
        std::ofstream o(dir + "/CELLParStruct.txt");
        if (!o) return gc::set_error(GC_ERR_INVALID, "cannot write CELLParStruct.txt");
        o << "nCell\n";
        put_int(o, d->nCell); o << "\n";
        int vb = 0, e0 = 0;
        for (int i = 0; i < d->nCell; ++i) {
            int tb = d->beginendtri[i].x, te = d->beginendtri[i].y, nt = te - tb + 1;
            int np = 0;
            while (vb + np < d->nCellPar && d->celltab[vb + np] == i) ++np;
            std::vector<float> key = {d->shearmod_para[e0], d->bendingmod_para[e0], d->hgarea_para[tb],
                                      d->hlarea_para[tb], d->hvol_para[tb]};
            int ty = -1;
            for (size_t k = 0; k < types.size(); ++k) if (types[k] == key) ty = (int)k;
            if (ty < 0) { types.push_back(key); ty = (int)types.size() - 1; }
            celltype[i] = ty;
            o << "Type nPar nTri nEdgeInfo (cell " << i << ")\n";
            put_int(o, ty + 1); put_int(o, np); put_int(o, nt); put_int(o, nt * 3 / 2); o << "\n";
            o << "x y z\n";
            for (int k = 0; k < np; ++k) {
                const tfloat3& p = d->coor[nsdpd + vb + k];
                put(o, p.x); put(o, p.y); put(o, p.z); o << "\n";
            }
            o << "v1 v2 v3 (1-based)\n";
            for (int t = tb; t <= te; ++t) {
                put_int(o, d->tri[t].x - vb); put_int(o, d->tri[t].y - vb); put_int(o, d->tri[t].z - vb);
                o << "\n";
            }
            o << "a b c d t1 t2\n";
            for (int k = 0; k < nt * 3 / 2; ++k) {
                const int6& E = d->edge[e0 + k];
                put_int(o, E.x - vb); put_int(o, E.y - vb); put_int(o, E.z - vb); put_int(o, E.k - vb);
                put_int(o, E.v - tb); put_int(o, E.w - tb); o << "\n";
            }
            vb += np;
            e0 += nt * 3 / 2;
        }
    }
    if (d->nCell > 0) {
        std::ofstream o(dir + "/CELLPhyPara.txt");
        if (!o) return gc::set_error(GC_ERR_INVALID, "cannot write CELLPhyPara.txt");
        o << "nType\n";
        put_int(o, (long)types.size()); o << "\n";
        const char* names[5] = {"ShearMod", "BendingMod", "HGarea", "HLarea", "HVol"};
        for (int k = 0; k < 5; ++k) {
            o << names[k] << "\n";
            for (auto& t : types) { put(o, t[k]); o << "\n"; }
        }
        o << "MaxCut\n"; put(o, c.MaxCut); o << "\n";
        o << "zeroLen\n"; put(o, c.zeroLen); o << "\n";
        o << "scaFac\n"; put(o, c.scaFac); o << "\n";
        o << "surEnergy\n"; put(o, c.surEnergy); o << "\n";
    }
    return GC_OK;
}

// -------------------------------------------------------------------------
// Tecplot frames of Particle::OutputPar (Particle.cu:681-1300) from an
// engine: same directories (Result/FluidCoor, FluidField, CellCoor,
// CellForce, CellGeo, CellNorDir under `dir`), same zone structure, slot
// order and number format (scientific, 5 digits, 20-column fields).
extern "C" int gc_EngineWriteTecplot(gc_engine* e, const char* dir_c, long stepdata, long stepcount,
                                     float timestep, long stepoutput, const tint3* tri,
                                     const tint2* beginendtri, const int* celltab)
{
    if (!e || !dir_c) return gc::set_error(GC_ERR_INVALID, "WriteTecplot: null argument");
    if (gc_EnsembleCount(e) != 1) return gc::set_error(GC_ERR_INVALID, "WriteTecplot: batched engine");
    gc_engine_info info;
    int st = gc_EngineInfo(e, &info);
    if (st) return st;
    const int N = info.nTotalPar, M = info.nCellPar, C = info.nCell;
    if (C > 0 && (!tri || !beginendtri || !celltab))
        return gc::set_error(GC_ERR_INVALID, "WriteTecplot: membrane topology needed");
    std::vector<tfloat3> coor(N), defF(M), aggF(M), nor, cen;
    std::vector<tfloat4> vr(N);
    std::vector<float> press(N), area(C), vol(C);
    gc_host_state hs;
    std::memset(&hs, 0, sizeof hs);
    hs.coor = coor.data(); hs.velrhop = vr.data(); hs.press = press.data();
    if (C > 0) { hs.area = area.data(); hs.vol = vol.data(); hs.defforce = defF.data(); hs.aggforce = aggF.data(); }
    if ((st = gc_EngineGetState(e, &hs))) return st;
    std::vector<int> sp(N);
    gc_host_cells hc;
    std::memset(&hc, 0, sizeof hc);
    hc.sortpart = sp.data();
    if ((st = gc_EngineGetCells(e, &hc))) return st;
    int ntri = 0;
    if (C > 0) {
        ntri = beginendtri[C - 1].y + 1;
        nor.resize(ntri); cen.resize(ntri);
        if ((st = gc_EngineGetTriangles(e, nor.data(), cen.data()))) return st;
    }
    // id ranges from the engine's parameter block
    const int nCellStart = N - M;
    int nB1 = 0, nB12 = 0;
    if ((st = gc_EngineRanges(e, &nB1, &nB12))) return st;
    const int nFB12 = nCellStart;
    std::string dir(dir_c);
    for (const char* sub : {"FluidCoor", "FluidField", "CellCoor", "CellForce", "CellGeo", "CellNorDir"})
        if ((st = mkdirs(dir + "/Result/" + sub))) return st;
    auto L = [](std::ostream& o) -> std::ostream& { return o << std::setiosflags(std::ios::left) << std::setw(20); };
    auto L15 = [](std::ostream& o) -> std::ostream& { return o << std::setiosflags(std::ios::left) << std::setw(15); };
    const std::string num = std::to_string(stepdata) + ".plt";
    // slot-order views
    auto at = [&](int slot) -> int { return sp[slot]; };
    {   // OutputCoor (Particle.cu:712-784)
        std::ofstream f(dir + "/Result/FluidCoor/" + num);
        L(f) << "Title=\"Particle Coordinate\"" << std::endl;
        L(f) << "Variables=\"x\",\"y\",\"z\",\"SP\"" << std::endl;
        f.precision(5);
        f.setf(std::ios::scientific, std::ios::floatfield);
        const int lo[4] = {0, nB1, nB12, nFB12}, hi[4] = {nB1, nB12, nFB12, N};
        for (int z = 0; z < 4; ++z) {
            if (hi[z] - lo[z] <= 0) continue;
            L(f) << "Zone" << std::endl;
            for (int i = 0; i < N; ++i) {
                int id = at(i);
                if (id >= lo[z] && id < hi[z]) {
                    L(f) << coor[id].x; L(f) << coor[id].y; L(f) << coor[id].z; L(f) << id;
                    f << std::endl;
                }
            }
        }
    }
    {   // OutputField (Particle.cu:906-941)
        std::ofstream f(dir + "/Result/FluidField/" + num);
        L(f) << "Title=\"Particle Fluid\"" << std::endl;
        L(f) << "Variables=\"x\",\"y\",\"z\",\"Den\",\"Press\",\"Vx\",\"Vy\",\"Vz\"" << std::endl;
        f.precision(5);
        f.setf(std::ios::scientific, std::ios::floatfield);
        for (int i = 0; i < N && nFB12 > 0; ++i) {
            int id = at(i);
            if (id >= 0 && id < nFB12) {
                L(f) << coor[id].x; L(f) << coor[id].y; L(f) << coor[id].z; L(f) << vr[id].w;
                L(f) << press[id]; L(f) << vr[id].x; L(f) << vr[id].y; L(f) << vr[id].z;
                f << std::endl;
            }
        }
    }
    if (C == 0) return GC_OK;
    // per-cell vertex ranges (BeginEndCoor, Particle.cu:223-225)
    std::vector<int> vb(C, 0), ve(C, -1);
    for (int v = 0, cprev = -1; v < M; ++v) {
        int c = celltab[v];
        if (c != cprev) { vb[c] = v; cprev = c; }
        ve[c] = v;
    }
    std::vector<int> pos(M, 0);
    const double t_out = (double)stepdata * timestep * stepoutput;
    {   // OutPutCellCoor (Particle.cu:943-1008)
        std::ofstream f(dir + "/Result/CellCoor/" + num);
        L(f) << "Title=\"Particle Coordinate\"" << std::endl;
        L(f) << "Variables=\"x\",\"y\",\"z\"" << std::endl;
        f.precision(5);
        f.setf(std::ios::scientific, std::ios::floatfield);
        for (int c = 0; c < C; ++c) {
            int nt = beginendtri[c].y - beginendtri[c].x + 1;
            f << "Zone N=" << (ve[c] - vb[c] + 1) << ", E=" << nt << ", F=FEPOINT, ET=TRIANGLE" << std::endl;
            int m = 1;
            for (int j = 0; j < N; ++j) {
                int k = at(j) - nCellStart;
                if (k >= vb[c] && k <= ve[c]) {
                    pos[k] = m++;
                    int id = at(j);
                    L(f) << coor[id].x; L(f) << coor[id].y; L(f) << coor[id].z; f << std::endl;
                }
            }
            for (int t = beginendtri[c].x; t <= beginendtri[c].y; ++t) {
                L(f) << pos[tri[t].x - 1]; L(f) << pos[tri[t].y - 1]; L(f) << pos[tri[t].z - 1];
                f << std::endl;
            }
        }
    }
    {   // OutPutCellForce (Particle.cu:1010-1157) + TF_Cell<i>.plt series
        std::ofstream f(dir + "/Result/CellForce/" + num);
        L(f) << "Title=\"Particle Force\"" << std::endl;
        L(f) << "Variables=\"x\",\"y\",\"z\", \"DFx\",\"DFy\",\"DFz\",\"AFx\",\"AFy\",\"AFz\"" << std::endl;
        f.precision(5);
        f.setf(std::ios::scientific, std::ios::floatfield);
        for (int c = 0; c < C; ++c) {
            int nt = beginendtri[c].y - beginendtri[c].x + 1;
            int np = ve[c] - vb[c] + 1;
            f << "Zone N=" << np << ", E=" << nt << ", F=FEPOINT, ET=TRIANGLE" << std::endl;
            int m = 1;
            double td[3] = {0, 0, 0}, ta[3] = {0, 0, 0};
            for (int j = 0; j < N; ++j) {
                int k = at(j) - nCellStart;
                if (k >= vb[c] && k <= ve[c]) {
                    pos[k] = m++;
                    int id = at(j);
                    L(f) << coor[id].x; L(f) << coor[id].y; L(f) << coor[id].z;
                    L(f) << defF[k].x; L(f) << defF[k].y; L(f) << defF[k].z;
                    L(f) << aggF[k].x; L(f) << aggF[k].y; L(f) << aggF[k].z;
                    f << std::endl;
                    td[0] += defF[k].x; td[1] += defF[k].y; td[2] += defF[k].z;
                    ta[0] += aggF[k].x; ta[1] += aggF[k].y; ta[2] += aggF[k].z;
                }
            }
            for (int t = beginendtri[c].x; t <= beginendtri[c].y; ++t) {
                L(f) << pos[tri[t].x - 1]; L(f) << pos[tri[t].y - 1]; L(f) << pos[tri[t].z - 1];
                f << std::endl;
            }
            std::ofstream g;
            std::string tf = dir + "/Result/CellForce/TF_Cell" + std::to_string(c) + ".plt";
            if (stepcount == 0) g.open(tf); else g.open(tf, std::ios::app | std::ios::out);
            if (!g) return gc::set_error(GC_ERR_INVALID, "cannot write " + tf);
            g.precision(5);
            g.setf(std::ios::scientific, std::ios::floatfield);
            L15(g) << t_out;
            for (int d = 0; d < 3; ++d) L15(g) << (float)(td[d] / np);
            for (int d = 0; d < 3; ++d) L15(g) << (float)(ta[d] / np);
            g << std::endl;
        }
    }
    {   // OutPutCellGeo (Particle.cu:1159-1246): |mean vertex velocity|, area, volume
        for (int c = 0; c < C; ++c) {
            double v[3] = {0, 0, 0};
            int np = ve[c] - vb[c] + 1;
            for (int k = vb[c]; k <= ve[c]; ++k) {
                const tfloat4& q = vr[nCellStart + k];
                v[0] += q.x; v[1] += q.y; v[2] += q.z;
            }
            float mag = (float)std::sqrt((v[0] * v[0] + v[1] * v[1] + v[2] * v[2])) / np;
            std::ofstream g;
            std::string gf = dir + "/Result/CellGeo/Geo_Cell" + std::to_string(c) + ".plt";
            if (stepcount == 0) g.open(gf); else g.open(gf, std::ios::app | std::ios::out);
            if (!g) return gc::set_error(GC_ERR_INVALID, "cannot write " + gf);
            g.precision(5);
            g.setf(std::ios::scientific, std::ios::floatfield);
            L15(g) << t_out; L15(g) << mag; L15(g) << area[c]; L15(g) << vol[c];
            g << std::endl;
        }
    }
    {   // OutPutTriNorDir (Particle.cu:1248-1300)
        std::ofstream f(dir + "/Result/CellNorDir/" + num);
        L(f) << "Title=\"Triangle information\"" << std::endl;
        L(f) << "Variables=\"Cenx\",\"Ceny\",\"Cenz\",\"Norx\",\"Nory\",\"Norz\",\"ID\"" << std::endl;
        f.precision(5);
        f.setf(std::ios::scientific, std::ios::floatfield);
        for (int c = 0; c < C; ++c) {
            L(f) << "Zone" << std::endl;
            for (int t = beginendtri[c].x; t <= beginendtri[c].y; ++t) {
                L(f) << cen[t].x; L(f) << cen[t].y; L(f) << cen[t].z;
                L(f) << nor[t].x; L(f) << nor[t].y; L(f) << nor[t].z; L(f) << t;
                f << std::endl;
            }
        }
    }
    return GC_OK;
}

// -------------------------------------------------------------------------
// Checkpoint / restart.  The reference has none (its outputs carry 5
// significant digits).  A checkpoint holds the engine's dynamic state in id
// order at full precision: FP64 positions, Velrhop, VelHalf, TotalForce,
// nThroBound, the step and the noise counter.  Restoring it into an engine
// created from the same case resumes bit for bit: the slot order is the
// (cube, id) order of the positions, which the restore rebuilds.
namespace {
const char kMagic[8] = {'G', 'C', 'E', 'L', 'L', 'C', 'K', '1'};
struct CkHeader {
    char magic[8];
    int64_t n, nm, ncell, step;
    uint64_t counter;
};
}  // namespace

extern "C" int gc_EngineSave(gc_engine* e, const char* path)
{
    if (!e || !path) return gc::set_error(GC_ERR_INVALID, "EngineSave: null argument");
    if (gc_EnsembleCount(e) != 1) return gc::set_error(GC_ERR_INVALID, "EngineSave: batched engine");
    gc_engine_info info;
    int st = gc_EngineInfo(e, &info);
    if (st) return st;
    const size_t N = (size_t)info.nTotalPar;
    std::vector<double> pos(3 * N);
    std::vector<tfloat4> vr(N);
    std::vector<tfloat3> vh(N), f(N);
    std::vector<tint3> nt(N);
    gc_host_state hs;
    std::memset(&hs, 0, sizeof hs);
    hs.pos = pos.data(); hs.velrhop = vr.data(); hs.velhalf = vh.data(); hs.force = f.data();
    hs.nthrobound = nt.data();
    if ((st = gc_EngineGetState(e, &hs))) return st;
    CkHeader h;
    std::memcpy(h.magic, kMagic, 8);
    h.n = (int64_t)N; h.nm = info.nCellPar; h.ncell = info.nCell; h.step = info.step;
    h.counter = info.noise_counter;
    std::string tmp = std::string(path) + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return gc::set_error(GC_ERR_INVALID, std::string("cannot write ") + tmp);
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1 &&
              std::fwrite(pos.data(), sizeof(double), 3 * N, fp) == 3 * N &&
              std::fwrite(vr.data(), sizeof(tfloat4), N, fp) == N &&
              std::fwrite(vh.data(), sizeof(tfloat3), N, fp) == N &&
              std::fwrite(f.data(), sizeof(tfloat3), N, fp) == N &&
              std::fwrite(nt.data(), sizeof(tint3), N, fp) == N;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0)
        return gc::set_error(GC_ERR_INVALID, std::string("write failed: ") + path);
    return GC_OK;
}

extern "C" int gc_EngineLoad(gc_engine* e, const char* path)
{
    if (!e || !path) return gc::set_error(GC_ERR_INVALID, "EngineLoad: null argument");
    if (gc_EnsembleCount(e) != 1) return gc::set_error(GC_ERR_INVALID, "EngineLoad: batched engine");
    gc_engine_info info;
    int st = gc_EngineInfo(e, &info);
    if (st) return st;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return gc::set_error(GC_ERR_INVALID, std::string("cannot open ") + path);
    CkHeader h;
    if (std::fread(&h, sizeof h, 1, fp) != 1 || std::memcmp(h.magic, kMagic, 8) != 0) {
        std::fclose(fp);
        return gc::set_error(GC_ERR_INVALID, std::string("not a checkpoint: ") + path);
    }
    if (h.n != info.nTotalPar || h.nm != info.nCellPar || h.ncell != info.nCell) {
        std::fclose(fp);
        return gc::set_error(GC_ERR_INVALID, "checkpoint was written by a different case");
    }
    const size_t N = (size_t)h.n;
    std::vector<double> pos(3 * N);
    std::vector<tfloat4> vr(N);
    std::vector<tfloat3> vh(N), f(N);
    std::vector<tint3> nt(N);
    bool ok = std::fread(pos.data(), sizeof(double), 3 * N, fp) == 3 * N &&
              std::fread(vr.data(), sizeof(tfloat4), N, fp) == N &&
              std::fread(vh.data(), sizeof(tfloat3), N, fp) == N &&
              std::fread(f.data(), sizeof(tfloat3), N, fp) == N &&
              std::fread(nt.data(), sizeof(tint3), N, fp) == N;
    std::fclose(fp);
    if (!ok) return gc::set_error(GC_ERR_INVALID, std::string("truncated checkpoint: ") + path);
    gc_host_state hs;
    std::memset(&hs, 0, sizeof hs);
    hs.pos = pos.data(); hs.velrhop = vr.data(); hs.velhalf = vh.data(); hs.force = f.data();
    hs.nthrobound = nt.data();
    return gc_EngineSetState(e, &hs, (long)h.step, h.counter);
}
