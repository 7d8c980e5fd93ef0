// output_host.cpp -- solution files on the host side of the output path:
// the .hsln reader (the layout Writer::write_soln writes, src/Writer.cpp:22-39,
// read back as interface/Export.py:14-32 does) for restarts, and a VTU writer
// that produces what interface/Export.py:34-81 produces with the vtk package
// (absent here): every cell of the mesh, ghost line cells included, with
// Density, Pressure and a 3-component Velocity as cell data.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace hcfd;

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(FILE* x) : f(x) {}
    ~File() {
        if (f) fclose(f);
    }
};

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

// One appended array of a VTU file: the XML attributes and its bytes.
struct Appended {
    std::string type, name;
    int ncomp = 1;
    std::vector<unsigned char> bytes;
};

template <typename T>
void put(Appended& a, const T* p, size_t n) {
    const size_t at = a.bytes.size();
    a.bytes.resize(at + n * sizeof(T));
    if (n) std::memcpy(a.bytes.data() + at, p, n * sizeof(T));
}

}  // namespace

extern "C" {

// .hsln reader: int32 Nc, nvars (= 4), nchars; nchars bytes of NUL-separated
// names; float64 [nvars][Nc].  Q == NULL only reports Nc.
int hcfd_read_hsln(const char* path, int* Nc, double* const Q[4]) {
    return guard([&] {
        if (!path || !Nc) fail(HCFD_ERR_ARG, "bad argument");
        const std::string p(path);
        File fh(fopen(path, "rb"));
        if (!fh.f) fail(HCFD_ERR_IO, p + ": cannot open");
        int hdr[3];
        if (fread(hdr, sizeof(int), 3, fh.f) != 3) fail(HCFD_ERR_IO, p + ": truncated .hsln header");
        if (hdr[0] <= 0 || hdr[1] != 4 || hdr[2] <= 0 || hdr[2] > 4096)
            fail(HCFD_ERR_IO, p + ": not a 2-D .hsln (Nc, nvars, nchars = " + std::to_string(hdr[0]) + ", " +
                                  std::to_string(hdr[1]) + ", " + std::to_string(hdr[2]) + ")");
        if (fseek(fh.f, 0, SEEK_END) != 0) fail(HCFD_ERR_IO, p + ": cannot seek");
        const long size = ftell(fh.f);
        const long want = 12L + hdr[2] + 8L * 4 * hdr[0];
        if (size != want)
            fail(HCFD_ERR_IO, p + ": size " + std::to_string(size) + " B, the header implies " + std::to_string(want));
        *Nc = hdr[0];
        if (!Q) return;
        if (fseek(fh.f, 12L + hdr[2], SEEK_SET) != 0) fail(HCFD_ERR_IO, p + ": cannot seek");
        for (int v = 0; v < 4; v++) {
            if (!Q[v]) fail(HCFD_ERR_ARG, "null array");
            if (fread(Q[v], sizeof(double), (size_t)hdr[0], fh.f) != (size_t)hdr[0])
                fail(HCFD_ERR_IO, p + ": truncated .hsln data");
        }
    });
}

// VTU (XML UnstructuredGrid, raw appended binary, UInt64 block headers).
// Cell types from the node count of each cell: 2 line, 3 triangle, 4 quad.
int hcfd_write_vtu(const char* path, const hcfd_mesh_view* m, const double* const Q[4]) {
    return guard([&] {
        if (!path || !m || !Q) fail(HCFD_ERR_ARG, "bad argument");
        if (m->dim != 2) fail(HCFD_ERR_ARG, "VTU writer: 2-D meshes only");
        if (m->Nc <= 0 || m->Nn <= 0 || !m->nodes || !m->cn_off || !m->cn_con) fail(HCFD_ERR_ARG, "bad mesh");
        for (int v = 0; v < 4; v++)
            if (!Q[v]) fail(HCFD_ERR_ARG, "null array");
        const int Nc = m->Nc, Nn = m->Nn;
        std::vector<Appended> arr(7);
        arr[0] = {"Float64", "Points", 3, {}};
        std::vector<double> xyz(3 * (size_t)Nn);
        for (int k = 0; k < Nn; k++) {
            xyz[3 * k] = m->nodes[2 * k];
            xyz[3 * k + 1] = m->nodes[2 * k + 1];
            xyz[3 * k + 2] = 0.0;
        }
        put(arr[0], xyz.data(), xyz.size());
        if (m->cn_off[0] != 0) fail(HCFD_ERR_MESH, "cn_off[0] != 0");
        std::vector<unsigned char> types(Nc);
        for (int i = 0; i < Nc; i++) {
            const int n = m->cn_off[i + 1] - m->cn_off[i];
            if (n < 2 || n > 4) fail(HCFD_ERR_MESH, "cell " + std::to_string(i) + " has " + std::to_string(n) + " nodes");
            types[i] = n == 2 ? 3 : n == 3 ? 5 : 9;  // VTK_LINE, VTK_TRIANGLE, VTK_QUAD
        }
        const int ncon = m->cn_off[Nc];
        for (int k = 0; k < ncon; k++)
            if (m->cn_con[k] < 0 || m->cn_con[k] >= Nn) fail(HCFD_ERR_MESH, "cn_con entry out of range");
        arr[1] = {"Int32", "connectivity", 1, {}};
        put(arr[1], m->cn_con, (size_t)ncon);
        arr[2] = {"Int32", "offsets", 1, {}};
        put(arr[2], m->cn_off + 1, (size_t)Nc);
        arr[3] = {"UInt8", "types", 1, {}};
        put(arr[3], types.data(), types.size());
        arr[4] = {"Float64", "Density", 1, {}};
        put(arr[4], Q[0], (size_t)Nc);
        arr[5] = {"Float64", "Pressure", 1, {}};
        put(arr[5], Q[3], (size_t)Nc);
        arr[6] = {"Float64", "Velocity", 3, {}};
        std::vector<double> vel(3 * (size_t)Nc);
        for (int i = 0; i < Nc; i++) {
            vel[3 * i] = Q[1][i];
            vel[3 * i + 1] = Q[2][i];
            vel[3 * i + 2] = 0.0;
        }
        put(arr[6], vel.data(), vel.size());

        std::vector<uint64_t> off(arr.size());
        uint64_t at = 0;
        for (size_t k = 0; k < arr.size(); k++) {
            off[k] = at;
            at += sizeof(uint64_t) + arr[k].bytes.size();
        }
        auto tag = [&](size_t k) {
            std::string s = "        <DataArray type=\"" + arr[k].type + "\" Name=\"" + arr[k].name + "\"";
            if (arr[k].ncomp > 1) s += " NumberOfComponents=\"" + std::to_string(arr[k].ncomp) + "\"";
            return s + " format=\"appended\" offset=\"" + std::to_string(off[k]) + "\"/>\n";
        };
        std::string x = "<?xml version=\"1.0\"?>\n"
                        "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                        "header_type=\"UInt64\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" +
                        std::to_string(Nn) + "\" NumberOfCells=\"" + std::to_string(Nc) + "\">\n      <Points>\n" +
                        tag(0) + "      </Points>\n      <Cells>\n" + tag(1) + tag(2) + tag(3) +
                        "      </Cells>\n      <CellData Scalars=\"Density\" Vectors=\"Velocity\">\n" + tag(4) +
                        tag(5) + tag(6) +
                        "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n_";
        const std::string p(path);
        File fh(fopen(path, "wb"));
        if (!fh.f) fail(HCFD_ERR_IO, p + ": cannot open for writing");
        bool ok = fwrite(x.data(), 1, x.size(), fh.f) == x.size();
        for (const auto& a : arr) {
            const uint64_t n = a.bytes.size();
            ok = ok && fwrite(&n, sizeof n, 1, fh.f) == 1;
            ok = ok && (n == 0 || fwrite(a.bytes.data(), 1, n, fh.f) == n);
        }
        const char* tail = "\n  </AppendedData>\n</VTKFile>\n";
        ok = ok && fwrite(tail, 1, strlen(tail), fh.f) == strlen(tail);
        if (!ok) fail(HCFD_ERR_IO, p + ": write failed");
    });
}

}  // extern "C"
