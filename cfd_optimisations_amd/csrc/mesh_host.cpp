// mesh_host.cpp -- host mesh services behind the C ABI: .hmsh reader/writer
// (the format of src/Mesh.cpp:11-45 and interface/Mesh.py:441-463), the .hsln
// writer (src/Writer.cpp:22-39) and seeded mesh generators.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh_struct.hpp"

namespace hcfd {

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(FILE* x) : f(x) {}
    ~File() {
        if (f) fclose(f);
    }
};

// Bytes left in the file after the current position.
size_t remaining(FILE* f) {
    const long pos = ftell(f);
    if (pos < 0 || fseek(f, 0, SEEK_END) != 0) return 0;
    const long end = ftell(f);
    fseek(f, pos, SEEK_SET);
    return end > pos ? (size_t)(end - pos) : 0;
}

// Reads n elements; a count the rest of the file cannot hold is reported as a
// truncated file before anything is allocated for it.
template <typename T>
void rd(FILE* f, std::vector<T>& v, size_t n, const char* what, const std::string& path) {
    if (n > remaining(f) / sizeof(T)) fail(HCFD_ERR_IO, path + ": truncated .hmsh while reading " + what);
    v.resize(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n)
        fail(HCFD_ERR_IO, path + ": truncated .hmsh while reading " + what);
}

// A CSR offset array: starts at 0 and never decreases.
void check_offsets(const std::vector<int>& off, const char* what, const std::string& path) {
    if (off.empty() || off[0] != 0) fail(HCFD_ERR_IO, path + ": " + what + " does not start at 0");
    for (size_t i = 1; i < off.size(); i++)
        if (off[i] < off[i - 1]) fail(HCFD_ERR_IO, path + ": " + what + " decreases at entry " + std::to_string(i));
}

template <typename T>
void wr(FILE* f, const T* p, size_t n, const std::string& path) {
    if (n && fwrite(p, sizeof(T), n, f) != n) fail(HCFD_ERR_IO, path + ": write failed");
}

// SplitMix64 (Steele, Lea, Flood 2014): portable, seedable, and defined the
// same way on every platform, unlike std::uniform_*_distribution.
struct SplitMix64 {
    uint64_t x;
    explicit SplitMix64(uint64_t s) : x(s) {}
    uint64_t next() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void build_cc(hcfd_mesh& m) {
    std::vector<int> deg(m.Nc + 1, 0);
    for (int f = 0; f < m.Nf; f++) {
        deg[m.fc_con[2 * f] + 1]++;
        deg[m.fc_con[2 * f + 1] + 1]++;
    }
    m.cc_off.assign(m.Nc + 1, 0);
    for (int i = 0; i < m.Nc; i++) m.cc_off[i + 1] = m.cc_off[i] + deg[i + 1];
    m.cc_con.assign(m.cc_off[m.Nc], 0);
    std::vector<int> cur(m.cc_off.begin(), m.cc_off.end() - 1);
    for (int f = 0; f < m.Nf; f++) {
        int a = m.fc_con[2 * f], b = m.fc_con[2 * f + 1];
        m.cc_con[cur[a]++] = b;
        m.cc_con[cur[b]++] = a;
    }
    for (int i = 0; i < m.Nc; i++) std::sort(m.cc_con.begin() + m.cc_off[i], m.cc_con.begin() + m.cc_off[i + 1]);
}

// Structured nx*ny quads over a node map, each quad split along (i,j)-(i+1,j+1)
// into two triangles; one ghost line cell per boundary edge.  Colours:
// 0 interior, 1 bottom, 2 right, 3 top, 4 left.  Face order follows the
// reference importer (interface/Mesh.py:321-378): interior faces sorted by
// (lower cell, higher cell) with CL the lower id, then boundary faces by
// colour with the ghost on the right.
template <typename NodeFn>
std::unique_ptr<hcfd_mesh> structured_triangles(int nx, int ny, NodeFn node_xy, double jitter, uint64_t seed,
                                                int shuffle, const char* const names[4]) {
    if (nx < 1 || ny < 1 || (long long)nx * ny > 200000000LL) fail(HCFD_ERR_ARG, "bad grid size");
    if (!(jitter >= 0.0 && jitter < 0.5)) fail(HCFD_ERR_ARG, "jitter must be in [0, 0.5)");
    auto m = std::make_unique<hcfd_mesh>();
    const int NX = nx + 1, NY = ny + 1;
    m->Nn = NX * NY;
    m->nodes.resize(2 * (size_t)m->Nn);
    SplitMix64 rng(seed);
    // Interior nodes move by up to +-jitter grid spacings in the LOGICAL (i, j)
    // coordinates before the node map; the maps have a positive Jacobian, so the
    // jittered cells keep their orientation even where the grid is sheared.
    for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
            double s = i, t = j;
            if (i > 0 && i < nx && j > 0 && j < ny && jitter > 0) {
                s += (2.0 * rng.uniform() - 1.0) * jitter;
                t += (2.0 * rng.uniform() - 1.0) * jitter;
            }
            double x, y;
            node_xy(s, t, x, y);
            m->nodes[2 * ((size_t)j * NX + i)] = x;
            m->nodes[2 * ((size_t)j * NX + i) + 1] = y;
        }
    auto nid = [&](int i, int j) { return j * NX + i; };
    const int Nint = 2 * nx * ny;
    const int Nghost = 2 * (nx + ny);
    m->Nc = Nint + Nghost;
    // cells: quad q = j*nx+i -> T0 = 2q (n00,n10,n11), T1 = 2q+1 (n00,n11,n01)
    m->cn_off.resize(m->Nc + 1);
    m->cn_con.reserve(3 * (size_t)Nint + 2 * (size_t)Nghost);
    m->c_ctypes.resize(m->Nc);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int q = j * nx + i;
            m->cn_off[2 * q] = (int)m->cn_con.size();
            m->cn_con.insert(m->cn_con.end(), {nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)});
            m->cn_off[2 * q + 1] = (int)m->cn_con.size();
            m->cn_con.insert(m->cn_con.end(), {nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)});
            m->c_ctypes[2 * q] = m->c_ctypes[2 * q + 1] = 2;
        }
    // faces
    struct F {
        int cl, cr, n1, n2;
    };
    std::vector<F> inner;
    inner.reserve(3 * (size_t)nx * ny);
    std::vector<F> bc[4];
    int g = Nint;
    std::vector<std::pair<int, int>> ghost_nodes;  // per ghost (in creation order)
    auto tri = [&](int i, int j, int t) { return 2 * (j * nx + i) + t; };
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            // diagonal inside the quad
            inner.push_back({tri(i, j, 0), tri(i, j, 1), nid(i, j), nid(i + 1, j + 1)});
            // right edge (x = i+1) between T0 of (i,j) and T1 of (i+1,j)
            if (i + 1 < nx) inner.push_back({tri(i, j, 0), tri(i + 1, j, 1), nid(i + 1, j), nid(i + 1, j + 1)});
            // top edge (y = j+1) between T1 of (i,j) and T0 of (i,j+1)
            if (j + 1 < ny) inner.push_back({tri(i, j, 1), tri(i, j + 1, 0), nid(i, j + 1), nid(i + 1, j + 1)});
        }
    for (int i = 0; i < nx; i++) bc[0].push_back({tri(i, 0, 0), -1, nid(i, 0), nid(i + 1, 0)});             // bottom
    for (int j = 0; j < ny; j++) bc[1].push_back({tri(nx - 1, j, 0), -1, nid(nx, j), nid(nx, j + 1)});     // right
    for (int i = 0; i < nx; i++) bc[2].push_back({tri(i, ny - 1, 1), -1, nid(i, ny), nid(i + 1, ny)});     // top
    for (int j = 0; j < ny; j++) bc[3].push_back({tri(0, j, 1), -1, nid(0, j), nid(0, j + 1)});            // left
    for (int s = 0; s < 4; s++)
        for (auto& f : bc[s]) {
            f.cr = g++;
            ghost_nodes.push_back({f.n1, f.n2});
        }
    for (int k = 0; k < Nghost; k++) {
        m->cn_off[Nint + k] = (int)m->cn_con.size();
        m->cn_con.push_back(ghost_nodes[k].first);
        m->cn_con.push_back(ghost_nodes[k].second);
        m->c_ctypes[Nint + k] = 1;
    }
    m->cn_off[m->Nc] = (int)m->cn_con.size();

    // optional random relabelling of all cells (Fisher-Yates with the same stream)
    std::vector<int> perm(m->Nc);
    for (int i = 0; i < m->Nc; i++) perm[i] = i;
    if (shuffle) {
        for (int i = m->Nc - 1; i > 0; i--) {
            int k = (int)(rng.next() % (uint64_t)(i + 1));
            std::swap(perm[i], perm[k]);
        }
        // perm[new] = old; relabel cell arrays
        std::vector<int> new_of_old(m->Nc);
        for (int i = 0; i < m->Nc; i++) new_of_old[perm[i]] = i;
        std::vector<int> off(m->Nc + 1, 0), con;
        std::vector<int> ct(m->Nc);
        con.reserve(m->cn_con.size());
        for (int n = 0; n < m->Nc; n++) {
            int o = perm[n];
            off[n] = (int)con.size();
            con.insert(con.end(), m->cn_con.begin() + m->cn_off[o], m->cn_con.begin() + m->cn_off[o + 1]);
            ct[n] = m->c_ctypes[o];
        }
        off[m->Nc] = (int)con.size();
        m->cn_off.swap(off);
        m->cn_con.swap(con);
        m->c_ctypes.swap(ct);
        for (auto& f : inner) {
            f.cl = new_of_old[f.cl];
            f.cr = new_of_old[f.cr];
        }
        for (int s = 0; s < 4; s++)
            for (auto& f : bc[s]) {
                f.cl = new_of_old[f.cl];
                f.cr = new_of_old[f.cr];
            }
    }
    for (auto& f : inner)
        if (f.cl > f.cr) std::swap(f.cl, f.cr);
    std::sort(inner.begin(), inner.end(), [](const F& a, const F& b) { return a.cl != b.cl ? a.cl < b.cl : a.cr < b.cr; });
    for (int s = 0; s < 4; s++)
        std::stable_sort(bc[s].begin(), bc[s].end(), [](const F& a, const F& b) { return a.cl < b.cl; });
    m->Ncol = 5;
    m->NBCcol = 4;
    m->coltag = {0, 1, 2, 3, 4};
    m->coloff.assign(6, 0);
    m->coloff[1] = (int)inner.size();
    for (int s = 0; s < 4; s++) m->coloff[s + 2] = m->coloff[s + 1] + (int)bc[s].size();
    m->Nf = m->coloff[5];
    m->fc_con.resize(2 * (size_t)m->Nf);
    m->fn_con.resize(2 * (size_t)m->Nf);
    size_t f = 0;
    auto put = [&](const F& x) {
        m->fc_con[2 * f] = x.cl;
        m->fc_con[2 * f + 1] = x.cr;
        m->fn_con[2 * f] = x.n1;
        m->fn_con[2 * f + 1] = x.n2;
        f++;
    };
    for (auto& x : inner) put(x);
    for (int s = 0; s < 4; s++)
        for (auto& x : bc[s]) put(x);
    for (int s = 0; s < 4; s++) m->pnames.push_back(names[s]);
    build_cc(*m);
    return m;
}

}  // namespace
}  // namespace hcfd

using namespace hcfd;

namespace {
template <typename F>
int guard(F&& f) {
    try {
        f();
        return HCFD_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HCFD_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(std::string("unexpected error: ") + e.what());
        return HCFD_ERR_ARG;
    }
}
}  // namespace

extern "C" {

// Mesh::Mesh (src/Mesh.cpp:7-52) with every read checked.
int hcfd_mesh_load(const char* path, hcfd_mesh** out) {
    return guard([&] {
        if (!path || !out) fail(HCFD_ERR_ARG, "null argument");
        std::string p(path);
        File fh(fopen(path, "rb"));
        if (!fh.f) fail(HCFD_ERR_IO, p + ": cannot open");
        auto m = std::make_unique<hcfd_mesh>();
        std::vector<int> hdr;
        rd(fh.f, hdr, 8, "header", p);
        m->dim = hdr[0];
        m->Nn = hdr[1];
        m->Ncol = hdr[2];
        m->NBCcol = hdr[3];
        m->Nc = hdr[4];
        m->Nf = hdr[5];
        const int npn = std::max(1, hdr[6]);  // src/Mesh.cpp:19
        const int nch = hdr[7];
        // counts up to INT_MAX - 1, so that Nc + 1 and Ncol + 1 entries are representable
        if (m->dim != 2 || m->Nn <= 0 || m->Ncol <= 0 || m->NBCcol < 0 || m->NBCcol > m->Ncol || m->Nc <= 0 ||
            m->Nf <= 0 || nch < 0 || m->Nc == INT_MAX || m->Ncol == INT_MAX)
            fail(HCFD_ERR_IO, p + ": bad .hmsh header");
        std::vector<char> raw;
        rd(fh.f, raw, nch, "names", p);
        // Mesh::parse_pnames (src/Mesh.cpp:70-86): N_pnames NUL-terminated strings
        size_t pos = 0;
        for (int k = 0; k < npn && pos <= raw.size(); k++) {
            size_t e = pos;
            while (e < raw.size() && raw[e] != '\0') e++;
            m->pnames.emplace_back(raw.data() + pos, raw.data() + e);
            pos = e + 1;
        }
        rd(fh.f, m->coltag, m->Ncol, "coltag", p);
        rd(fh.f, m->coloff, (size_t)m->Ncol + 1, "coloff", p);
        rd(fh.f, m->nodes, 2 * (size_t)m->Nn, "nodes", p);
        rd(fh.f, m->fn_con, 2 * (size_t)m->Nf, "fn_con", p);
        rd(fh.f, m->fc_con, 2 * (size_t)m->Nf, "fc_con", p);
        check_offsets(m->coloff, "coloff", p);
        if (m->coloff[m->Ncol] != m->Nf) fail(HCFD_ERR_IO, p + ": coloff does not end at Nf");
        rd(fh.f, m->cn_off, (size_t)m->Nc + 1, "cn_off", p);
        check_offsets(m->cn_off, "cn_off", p);
        rd(fh.f, m->cn_con, m->cn_off[m->Nc], "cn_con", p);
        rd(fh.f, m->cc_off, (size_t)m->Nc + 1, "cc_off", p);
        check_offsets(m->cc_off, "cc_off", p);
        rd(fh.f, m->cc_con, m->cc_off[m->Nc], "cc_con", p);
        m->c_ctypes.resize(m->Nc);
        if (fread(m->c_ctypes.data(), sizeof(int), m->Nc, fh.f) != (size_t)m->Nc) m->c_ctypes.assign(m->Nc, 0);
        *out = m.release();
    });
}

int hcfd_mesh_save(const hcfd_mesh_view* v, const char* path) {
    return guard([&] {
        if (!v || !path) fail(HCFD_ERR_ARG, "null argument");
        std::string p(path);
        // connectivity of cells sharing a face, and physical names are not in a view
        hcfd_mesh m;
        m.Nc = v->Nc;
        m.Nf = v->Nf;
        m.fc_con.assign(v->fc_con, v->fc_con + 2 * (size_t)v->Nf);
        build_cc(m);
        File fh(fopen(path, "wb"));
        if (!fh.f) fail(HCFD_ERR_IO, p + ": cannot open for writing");
        // one name per BC colour (a view carries no names; Mesh::parse_pnames needs >= 1)
        std::string names;
        const int nn = std::max(1, v->NBCcol);
        for (int k = 0; k < nn; k++) names += "bc" + std::to_string(k) + '\0';
        int hdr[8] = {v->dim, v->Nn, v->Ncol, v->NBCcol, v->Nc, v->Nf, nn, (int)names.size()};
        wr(fh.f, hdr, 8, p);
        wr(fh.f, names.data(), names.size(), p);
        wr(fh.f, v->coltag, v->Ncol, p);
        wr(fh.f, v->coloff, v->Ncol + 1, p);
        wr(fh.f, v->nodes, 2 * (size_t)v->Nn, p);
        wr(fh.f, v->fn_con, 2 * (size_t)v->Nf, p);
        wr(fh.f, v->fc_con, 2 * (size_t)v->Nf, p);
        wr(fh.f, v->cn_off, v->Nc + 1, p);
        wr(fh.f, v->cn_con, v->cn_off[v->Nc], p);
        wr(fh.f, m.cc_off.data(), m.cc_off.size(), p);
        wr(fh.f, m.cc_con.data(), m.cc_con.size(), p);
        std::vector<int> ct(v->Nc);
        for (int i = 0; i < v->Nc; i++) ct[i] = (v->cn_off[i + 1] - v->cn_off[i] == 2) ? 1 : (v->cn_off[i + 1] - v->cn_off[i] == 3 ? 2 : 3);
        wr(fh.f, ct.data(), ct.size(), p);
    });
}

int hcfd_mesh_view_of(const hcfd_mesh* m, hcfd_mesh_view* o) {
    return guard([&] {
        if (!m || !o) fail(HCFD_ERR_ARG, "null argument");
        o->dim = m->dim;
        o->Nn = m->Nn;
        o->Ncol = m->Ncol;
        o->NBCcol = m->NBCcol;
        o->Nc = m->Nc;
        o->Nf = m->Nf;
        o->coltag = m->coltag.data();
        o->coloff = m->coloff.data();
        o->nodes = m->nodes.data();
        o->fn_con = m->fn_con.data();
        o->fc_con = m->fc_con.data();
        o->cn_off = m->cn_off.data();
        o->cn_con = m->cn_con.data();
    });
}

int hcfd_mesh_free(hcfd_mesh* m) {
    delete m;
    return HCFD_OK;
}

const char* hcfd_mesh_pname(const hcfd_mesh* m, int k) {
    if (!m || k < 0 || k >= (int)m->pnames.size()) return nullptr;
    return m->pnames[k].c_str();
}

int hcfd_meshgen_riemann(int nx, int ny, double jitter, unsigned long long seed, int shuffle, hcfd_mesh** out) {
    return guard([&] {
        if (!out) fail(HCFD_ERR_ARG, "null out");
        static const char* names[4] = {"bottom", "right", "top", "left"};
        auto fn = [&](double i, double j, double& x, double& y) {
            x = i / nx;
            y = j / ny;
        };
        *out = structured_triangles(nx, ny, fn, jitter, seed, shuffle, names).release();
    });
}

int hcfd_meshgen_ramp(int nx, int ny, double length, double height, double x_ramp, double angle_deg, double jitter,
                      unsigned long long seed, int shuffle, hcfd_mesh** out) {
    return guard([&] {
        if (!out) fail(HCFD_ERR_ARG, "null out");
        if (!(length > 0 && height > 0 && x_ramp >= 0 && x_ramp < length && angle_deg >= 0 && angle_deg < 60))
            fail(HCFD_ERR_ARG, "bad ramp geometry");
        const double ta = std::tan(angle_deg * M_PI / 180.0);
        if ((length - x_ramp) * ta >= 0.95 * height) fail(HCFD_ERR_ARG, "ramp closes the channel");
        static const char* names[4] = {"lower", "outlet", "upper", "inlet"};
        auto fn = [&](double i, double j, double& x, double& y) {
            x = length * i / nx;
            const double yb = x > x_ramp ? (x - x_ramp) * ta : 0.0;
            y = yb + (height - yb) * j / ny;
        };
        *out = structured_triangles(nx, ny, fn, jitter, seed, shuffle, names).release();
    });
}

// Writer::write_soln byte layout (src/Writer.cpp:27-35), with the caller's
// variable names (the Writer's varnames, src/Writer.cpp:13); NULL gives the
// names Euler passes (src/Base/Euler.cpp:18).
int hcfd_write_hsln(const char* path, int Nc, const double* const Q[4], const char* const varnames[4]) {
    return guard([&] {
        if (!path || !Q || Nc <= 0) fail(HCFD_ERR_ARG, "bad argument");
        static const char* const dflt[4] = {"Density", "Velocity_X", "Velocity_Y", "Pressure"};
        const char* const* names = varnames ? varnames : dflt;
        int nchars = 0;
        for (int v = 0; v < 4; v++) {
            if (!names[v]) fail(HCFD_ERR_ARG, "null variable name");
            nchars += (int)strlen(names[v]) + 1;
        }
        for (int v = 0; v < 4; v++)
            if (!Q[v]) fail(HCFD_ERR_ARG, "null array");
        std::string p(path);
        File fh(fopen(path, "wb"));
        if (!fh.f) fail(HCFD_ERR_IO, p + ": cannot open for writing");
        int hdr[3] = {Nc, 4, nchars};
        wr(fh.f, hdr, 3, p);
        for (int v = 0; v < 4; v++) wr(fh.f, names[v], strlen(names[v]) + 1, p);
        for (int v = 0; v < 4; v++) wr(fh.f, Q[v], Nc, p);
    });
}

}  // extern "C"
