// msh_import.cpp -- gmsh v4 ASCII .msh -> mesh, with the conventions of the
// reference's importer (interface/Mesh.py:21-127, 159-378):
//   * physical names: the first N-1 of $PhysicalNames (the last, the surface, is dropped);
//   * elements of physical curves become boundary line cells tagged with the
//     curve's physical tag; surface elements get tag 0 (interior);
//   * element blocks sorted by tag and merged per (tag, type);
//   * cells are neighbours when they share >= 2 nodes (an edge);
//   * cells ordered by reverse Cuthill-McKee, nodes by Cuthill-McKee of the
//     node graph (interface/Mesh.py:296-319);
//   * one face per neighbour pair, CL the lower cell id unless its tag is
//     higher (the ghost is always on the right), face colour = max tag,
//     faces grouped by colour (interface/Mesh.py:321-378).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "mesh_struct.hpp"

namespace hcfd {
namespace {

const int kNodesOfType[8] = {0, 2, 3, 4, 4, 8, 6, 5};  // gmsh element types 1..7

std::string section(const std::string& all, const std::string& name, const std::string& path) {
    const std::string s = "$" + name, e = "$End" + name;
    size_t a = all.find(s);
    size_t b = all.find(e);
    if (a == std::string::npos || b == std::string::npos || b < a)
        fail(HCFD_ERR_IO, path + ": missing section " + s);
    return all.substr(a + s.size(), b - a - s.size());
}

// Cuthill-McKee order of a graph in CSR form (BFS from a minimum-degree vertex
// of each component, siblings by increasing degree).
std::vector<int> cuthill_mckee(int n, const std::vector<int>& off, const std::vector<int>& adj) {
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0);
    std::vector<int> verts(n);
    for (int i = 0; i < n; i++) verts[i] = i;
    std::stable_sort(verts.begin(), verts.end(),
                     [&](int a, int b) { return off[a + 1] - off[a] < off[b + 1] - off[b]; });
    std::vector<int> kids;
    for (int s : verts) {
        if (seen[s]) continue;
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            int p = order[head++];
            kids.clear();
            for (int k = off[p]; k < off[p + 1]; k++)
                if (!seen[adj[k]]) {
                    seen[adj[k]] = 1;
                    kids.push_back(adj[k]);
                }
            std::stable_sort(kids.begin(), kids.end(),
                             [&](int a, int b) { return off[a + 1] - off[a] < off[b + 1] - off[b]; });
            order.insert(order.end(), kids.begin(), kids.end());
        }
    }
    return order;
}

std::unique_ptr<hcfd_mesh> import_msh(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail(HCFD_ERR_IO, path + ": cannot open");
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string all = ss.str();
    {
        std::istringstream fmt(section(all, "MeshFormat", path));
        double ver = 0;
        int ftype = -1;
        fmt >> ver >> ftype;
        if (ver < 4.0 || ver >= 5.0 || ftype != 0) fail(HCFD_ERR_IO, path + ": need gmsh v4 ASCII (.msh 4.x, file-type 0)");
    }
    auto m = std::make_unique<hcfd_mesh>();
    // physical names (interface/Mesh.py:48-55)
    {
        std::istringstream pn(section(all, "PhysicalNames", path));
        int n = 0;
        pn >> n;
        std::string line;
        std::getline(pn, line);
        for (int i = 0; i < n; i++) {
            std::getline(pn, line);
            size_t q1 = line.find('"'), q2 = line.rfind('"');
            if (q1 == std::string::npos || q2 <= q1) fail(HCFD_ERR_IO, path + ": bad $PhysicalNames line");
            if (i < n - 1) m->pnames.push_back(line.substr(q1 + 1, q2 - q1 - 1));
        }
    }
    // curve physical tags (interface/Mesh.py:57-71)
    std::unordered_map<int, int> curve_tag;
    {
        std::istringstream en(section(all, "Entities", path));
        long npts, ncur, nsurf, nvol;
        en >> npts >> ncur >> nsurf >> nvol;
        if (nvol != 0) fail(HCFD_ERR_IO, path + ": 3-D meshes are not supported");
        if (nsurf == 0) fail(HCFD_ERR_IO, path + ": no surface");
        std::string line;
        std::getline(en, line);
        for (long i = 0; i < npts; i++) std::getline(en, line);
        for (long i = 0; i < ncur; i++) {
            std::getline(en, line);
            std::istringstream cl(line);
            int tag, nphys;
            double bb[6];
            cl >> tag >> bb[0] >> bb[1] >> bb[2] >> bb[3] >> bb[4] >> bb[5] >> nphys;
            if (nphys > 1) fail(HCFD_ERR_IO, path + ": curve with more than one physical tag");
            int pt = 0;
            if (nphys == 1) cl >> pt;
            curve_tag[tag] = nphys == 1 ? pt : -1;
        }
    }
    // nodes (interface/Mesh.py:73-97); tags mapped to read order
    std::vector<double> xyz;
    std::unordered_map<long, int> node_index;
    {
        std::istringstream nd(section(all, "Nodes", path));
        long nblocks, nnodes, mintag, maxtag;
        nd >> nblocks >> nnodes >> mintag >> maxtag;
        if (!nd || nblocks < 0 || nnodes < 0 || nnodes > INT_MAX / 2) fail(HCFD_ERR_IO, path + ": bad $Nodes header");
        xyz.reserve(2 * (size_t)nnodes);
        for (long b = 0; b < nblocks; b++) {
            int edim, etag, para;
            long n;
            nd >> edim >> etag >> para >> n;
            if (!nd || n < 0 || n > nnodes) fail(HCFD_ERR_IO, path + ": bad $Nodes block header");
            if (para) fail(HCFD_ERR_IO, path + ": parametric nodes are not supported");
            std::vector<long> tags(n);
            for (long k = 0; k < n; k++) nd >> tags[k];
            for (long k = 0; k < n; k++) {
                double x, y, z;
                nd >> x >> y >> z;
                node_index[tags[k]] = (int)(xyz.size() / 2);
                xyz.push_back(x);
                xyz.push_back(y);
            }
        }
        if (!nd) fail(HCFD_ERR_IO, path + ": truncated $Nodes");
    }
    const int Nn = (int)(xyz.size() / 2);
    // element blocks (interface/Mesh.py:99-127), then sort/merge by (tag, type)
    struct Block {
        int ptag, type;
        std::vector<int> con;
    };
    std::vector<Block> blocks;
    {
        std::istringstream el(section(all, "Elements", path));
        long nblocks, nelem, mintag, maxtag;
        el >> nblocks >> nelem >> mintag >> maxtag;
        if (!el || nblocks < 0 || nelem < 0 || nelem > INT_MAX) fail(HCFD_ERR_IO, path + ": bad $Elements header");
        for (long b = 0; b < nblocks; b++) {
            int edim, etag, etype;
            long n;
            el >> edim >> etag >> etype >> n;
            if (!el || n < 0 || n > nelem) fail(HCFD_ERR_IO, path + ": bad $Elements block header");
            if (etype < 1 || etype > 7) fail(HCFD_ERR_IO, path + ": unsupported element type " + std::to_string(etype));
            const int nn = kNodesOfType[etype];
            Block blk;
            blk.type = etype;
            blk.ptag = 0;
            if (edim == 1) {
                auto it = curve_tag.find(etag);
                if (it == curve_tag.end() || it->second < 0) fail(HCFD_ERR_IO, path + ": line elements on an untagged curve");
                blk.ptag = it->second;
            } else if (edim != 2) {
                fail(HCFD_ERR_IO, path + ": elements of dimension " + std::to_string(edim) + " are not supported");
            }
            blk.con.reserve(n * nn);
            for (long k = 0; k < n; k++) {
                long et;
                el >> et;
                for (int q = 0; q < nn; q++) {
                    long t;
                    el >> t;
                    auto it = node_index.find(t);
                    if (it == node_index.end()) fail(HCFD_ERR_IO, path + ": element references an unknown node");
                    blk.con.push_back(it->second);
                }
            }
            if (!el) fail(HCFD_ERR_IO, path + ": truncated $Elements");
            blocks.push_back(std::move(blk));
        }
    }
    std::stable_sort(blocks.begin(), blocks.end(), [](const Block& a, const Block& b) { return a.ptag < b.ptag; });
    std::vector<Block> merged;
    for (auto& b : blocks) {
        bool done = false;
        for (auto& m2 : merged)
            if (m2.ptag == b.ptag && m2.type == b.type) {
                m2.con.insert(m2.con.end(), b.con.begin(), b.con.end());
                done = true;
                break;
            }
        if (!done) merged.push_back(std::move(b));
    }
    // cells in block order
    std::vector<int> c_off{0}, c_con, c_tag, c_type;
    for (auto& b : merged) {
        const int nn = kNodesOfType[b.type];
        for (size_t k = 0; k < b.con.size() / nn; k++) {
            for (int q = 0; q < nn; q++) c_con.push_back(b.con[k * nn + q]);
            c_off.push_back((int)c_con.size());
            c_tag.push_back(b.ptag);
            c_type.push_back(b.type);
        }
    }
    const int Nc = (int)c_tag.size();
    // cell adjacency through shared edges (>= 2 shared nodes)
    std::unordered_map<uint64_t, std::vector<int>> edge_cells;
    auto ekey = [](int a, int b) { return a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a; };
    for (int i = 0; i < Nc; i++) {
        const int n = c_off[i + 1] - c_off[i];
        const int corners = (c_type[i] == 1) ? 1 : (c_type[i] == 2 ? 3 : 4);
        for (int e = 0; e < corners; e++) {
            int a = c_con[c_off[i] + e], b = c_con[c_off[i] + (e + 1) % std::max(n, 2)];
            edge_cells[ekey(a, b)].push_back(i);
        }
    }
    std::vector<std::vector<int>> nbr(Nc);
    for (auto& kv : edge_cells) {
        auto& v = kv.second;
        for (size_t a = 0; a < v.size(); a++)
            for (size_t b = a + 1; b < v.size(); b++) {
                nbr[v[a]].push_back(v[b]);
                nbr[v[b]].push_back(v[a]);
            }
    }
    std::vector<int> off(Nc + 1, 0), adj;
    for (int i = 0; i < Nc; i++) {
        std::sort(nbr[i].begin(), nbr[i].end());
        nbr[i].erase(std::unique(nbr[i].begin(), nbr[i].end()), nbr[i].end());
        off[i + 1] = off[i] + (int)nbr[i].size();
        adj.insert(adj.end(), nbr[i].begin(), nbr[i].end());
    }
    // cells: reverse Cuthill-McKee
    std::vector<int> cm = cuthill_mckee(Nc, off, adj);
    std::vector<int> cnew(Nc);
    for (int r = 0; r < Nc; r++) cnew[cm[Nc - 1 - r]] = r;
    // nodes: Cuthill-McKee of the node graph (reversed RCM)
    std::vector<std::vector<int>> nn_adj(Nn);
    for (int i = 0; i < Nc; i++)
        for (int a = c_off[i]; a < c_off[i + 1]; a++)
            for (int b = c_off[i]; b < c_off[i + 1]; b++)
                if (a != b) nn_adj[c_con[a]].push_back(c_con[b]);
    std::vector<int> noff(Nn + 1, 0), nadj;
    for (int i = 0; i < Nn; i++) {
        std::sort(nn_adj[i].begin(), nn_adj[i].end());
        nn_adj[i].erase(std::unique(nn_adj[i].begin(), nn_adj[i].end()), nn_adj[i].end());
        noff[i + 1] = noff[i] + (int)nn_adj[i].size();
        nadj.insert(nadj.end(), nn_adj[i].begin(), nn_adj[i].end());
    }
    std::vector<int> ncm = cuthill_mckee(Nn, noff, nadj);
    std::vector<int> nnew(Nn);
    for (int r = 0; r < Nn; r++) nnew[ncm[r]] = r;

    m->Nn = Nn;
    m->Nc = Nc;
    m->nodes.resize(2 * (size_t)Nn);
    for (int i = 0; i < Nn; i++) {
        m->nodes[2 * nnew[i]] = xyz[2 * i];
        m->nodes[2 * nnew[i] + 1] = xyz[2 * i + 1];
    }
    std::vector<int> old_of_new(Nc);
    for (int i = 0; i < Nc; i++) old_of_new[cnew[i]] = i;
    m->cn_off.assign(Nc + 1, 0);
    m->c_ctypes.resize(Nc);
    std::vector<int> tag_new(Nc);
    for (int r = 0; r < Nc; r++) {
        int o = old_of_new[r];
        for (int a = c_off[o]; a < c_off[o + 1]; a++) m->cn_con.push_back(nnew[c_con[a]]);
        m->cn_off[r + 1] = (int)m->cn_con.size();
        m->c_ctypes[r] = c_type[o];
        tag_new[r] = c_tag[o];
    }
    // Colours by physical tag: 0 for the surface cells, 1..N-1 for the boundary
    // curves' tags, which must be exactly those values.  (The reference counts
    // colours by (tag, element type) block and drops the faces of any tag at or
    // past that count, interface/Mesh.py:362-372; with contiguous tags and one
    // element type per tag the two agree.)
    std::vector<int> tags;
    for (auto& b : merged) tags.push_back(b.ptag);
    std::sort(tags.begin(), tags.end());
    tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
    for (size_t k = 0; k < tags.size(); k++)
        if (tags[k] != (int)k)
            fail(HCFD_ERR_IO, path + ": the boundary curves' physical tags must be 1.." +
                                  std::to_string(tags.size() - 1) + " (tag " + std::to_string(tags[k]) +
                                  " found where " + std::to_string(k) + " was expected)");
    // faces: one per neighbour pair (i < j in the new numbering)
    const int ncol = (int)tags.size();
    struct F {
        int cl, cr, n1, n2, col;
    };
    std::vector<std::vector<F>> bycol(ncol);
    for (int r = 0; r < Nc; r++) {
        int o = old_of_new[r];
        for (int q : nbr[o]) {
            int s = cnew[q];
            if (s <= r) continue;
            // shared nodes, ascending new ids
            std::vector<int> sh;
            for (int a = c_off[o]; a < c_off[o + 1]; a++)
                for (int b = c_off[q]; b < c_off[q + 1]; b++)
                    if (c_con[a] == c_con[b]) sh.push_back(nnew[c_con[a]]);
            std::sort(sh.begin(), sh.end());
            sh.erase(std::unique(sh.begin(), sh.end()), sh.end());
            if (sh.size() < 2) continue;
            F f{r, s, sh[0], sh[1], std::max(tag_new[r], tag_new[s])};
            if (tag_new[r] > tag_new[s]) std::swap(f.cl, f.cr);
            if (f.col >= ncol) fail(HCFD_ERR_IO, path + ": face of an unknown colour");
            bycol[f.col].push_back(f);
        }
    }
    // the importer indexes colours by tag value (interface/Mesh.py:362-372)
    m->Ncol = ncol;
    m->NBCcol = ncol - 1;
    m->coltag.resize(ncol);
    m->coloff.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; c++) {
        m->coltag[c] = c;
        for (auto& f : bycol[c]) {
            m->fc_con.push_back(f.cl);
            m->fc_con.push_back(f.cr);
            m->fn_con.push_back(f.n1);
            m->fn_con.push_back(f.n2);
        }
        m->coloff[c + 1] = (int)(m->fc_con.size() / 2);
    }
    m->Nf = m->coloff[ncol];
    // cell -> cell CSR in the new numbering
    m->cc_off.assign(Nc + 1, 0);
    std::vector<std::vector<int>> cc(Nc);
    for (int f = 0; f < m->Nf; f++) {
        cc[m->fc_con[2 * f]].push_back(m->fc_con[2 * f + 1]);
        cc[m->fc_con[2 * f + 1]].push_back(m->fc_con[2 * f]);
    }
    for (int i = 0; i < Nc; i++) {
        std::sort(cc[i].begin(), cc[i].end());
        m->cc_off[i + 1] = m->cc_off[i] + (int)cc[i].size();
        m->cc_con.insert(m->cc_con.end(), cc[i].begin(), cc[i].end());
    }
    return m;
}

}  // namespace
}  // namespace hcfd

extern "C" int hcfd_mesh_import_msh(const char* path, hcfd_mesh** out) {
    try {
        if (!path || !out) hcfd::fail(HCFD_ERR_ARG, "null argument");
        *out = hcfd::import_msh(path).release();
        return HCFD_OK;
    } catch (const hcfd::Error& e) {
        hcfd::set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        hcfd::set_last_error("host allocation failed");
        return HCFD_ERR_NOMEM;
    } catch (const std::exception& e) {
        hcfd::set_last_error(std::string("unexpected error: ") + e.what());
        return HCFD_ERR_IO;
    }
}
