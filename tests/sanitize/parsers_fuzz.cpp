// parsers_fuzz.cpp -- drives the library's host-side file readers (built with
// the AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize.py)
// over valid files and seeded corruptions of them: truncations, byte flips and
// overwritten count fields.  Every call must return a status code (0 or an
// error); any out-of-bounds access, leak or undefined behaviour on the way is a
// sanitizer report, which fails the test.
//
//   parsers_fuzz <work dir> <mesh.hmsh> <solution.hsln> <mesh.msh>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hcfd.h"

namespace {

std::vector<unsigned char> slurp(const std::string& p) {
    std::vector<unsigned char> b;
    if (std::FILE* f = std::fopen(p.c_str(), "rb")) {
        int c;
        while ((c = std::fgetc(f)) != EOF) b.push_back((unsigned char)c);
        std::fclose(f);
    }
    return b;
}

void spill(const std::string& p, const std::vector<unsigned char>& b) {
    std::FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", p.c_str());
        std::exit(2);
    }
    if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

// splitmix64: the corruptions are the same on every run
unsigned long long g_state = 0x9e3779b97f4a7c15ull;
unsigned long long next() {
    unsigned long long z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int g_ok = 0, g_err = 0;

void tally(int rc) {
    if (rc == HCFD_OK) g_ok++;
    else g_err++;
}

// Each reader, on one file: success or a status code, and every mesh freed.
void try_hmsh(const std::string& p) {
    hcfd_mesh* m = nullptr;
    const int rc = hcfd_mesh_load(p.c_str(), &m);
    tally(rc);
    if (rc == HCFD_OK) {
        hcfd_mesh_view v;
        if (hcfd_mesh_view_of(m, &v) != HCFD_OK) std::abort();
        hcfd_mesh_free(m);
    }
}

void try_msh(const std::string& p) {
    hcfd_mesh* m = nullptr;
    const int rc = hcfd_mesh_import_msh(p.c_str(), &m);
    tally(rc);
    if (rc == HCFD_OK) hcfd_mesh_free(m);
}

void try_hsln(const std::string& p) {
    int n = 0;
    int rc = hcfd_read_hsln(p.c_str(), &n, nullptr);
    if (rc == HCFD_OK && n > 0 && n < (1 << 22)) {
        std::vector<double> q(4 * (size_t)n);
        double* Q[4] = {&q[0], &q[(size_t)n], &q[2 * (size_t)n], &q[3 * (size_t)n]};
        rc = hcfd_read_hsln(p.c_str(), &n, Q);
    }
    tally(rc);
}

// Truncations at `cuts` offsets, `flips` single-byte changes, and `words` 32-bit
// fields of the first `head` bytes overwritten with extreme counts.
template <typename F>
void corrupt(const std::string& dir, const std::string& tag, const std::vector<unsigned char>& good, F reader,
             int cuts, int flips, int words, size_t head) {
    const std::string p = dir + "/fuzz_" + tag;
    for (int k = 0; k < cuts; k++) {
        std::vector<unsigned char> b(good.begin(), good.begin() + (good.size() * k) / cuts);
        spill(p, b);
        reader(p);
    }
    for (int k = 0; k < flips; k++) {
        std::vector<unsigned char> b = good;
        b[next() % b.size()] ^= (unsigned char)(1u << (next() % 8));
        spill(p, b);
        reader(p);
    }
    const int extreme[] = {-1, 0, 1, 2, 3, 0x7fffffff, (int)0x80000000, 1 << 20, 1 << 28};
    const size_t span = std::min(head, good.size()) / 4;
    for (int k = 0; k < words && span > 0; k++) {
        std::vector<unsigned char> b = good;
        const size_t at = 4 * (next() % span);
        const int x = extreme[next() % (sizeof extreme / sizeof extreme[0])];
        std::memcpy(&b[at], &x, 4);
        spill(p, b);
        reader(p);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: parsers_fuzz <work dir> <mesh.hmsh> <solution.hsln> <mesh.msh>\n");
        return 2;
    }
    const std::string dir = argv[1];
    // valid inputs first: the shipped mesh, generated meshes through save and load,
    // the VTU writer on one of them
    try_hmsh(argv[2]);
    for (int shuffle = 0; shuffle < 2; shuffle++) {
        hcfd_mesh* m = nullptr;
        if (hcfd_meshgen_riemann(7, 5, 0.2, 3, shuffle, &m) != HCFD_OK) return 3;
        hcfd_mesh_view v;
        if (hcfd_mesh_view_of(m, &v) != HCFD_OK) return 3;
        const std::string p = dir + "/gen.hmsh";
        if (hcfd_mesh_save(&v, p.c_str()) != HCFD_OK) return 3;
        try_hmsh(p);
        std::vector<double> q(4 * (size_t)v.Nc, 1.0);
        const double* Q[4] = {&q[0], &q[(size_t)v.Nc], &q[2 * (size_t)v.Nc], &q[3 * (size_t)v.Nc]};
        if (hcfd_write_vtu((dir + "/gen.vtu").c_str(), &v, Q) != HCFD_OK) return 3;
        if (hcfd_write_hsln((dir + "/gen.hsln").c_str(), v.Nc, Q, nullptr) != HCFD_OK) return 3;
        try_hsln(dir + "/gen.hsln");
        hcfd_mesh_free(m);
    }
    if (g_err) {
        std::fprintf(stderr, "a valid input was refused: %s\n", hcfd_last_error());
        return 4;
    }
    try_hsln(argv[3]);
    try_msh(argv[4]);
    if (g_err) {
        std::fprintf(stderr, "a valid input was refused: %s\n", hcfd_last_error());
        return 4;
    }
    corrupt(dir, "hmsh", slurp(argv[2]), try_hmsh, 97, 300, 200, 4096);
    corrupt(dir, "hsln", slurp(argv[3]), try_hsln, 61, 200, 60, 256);
    corrupt(dir, "msh", slurp(argv[4]), try_msh, 97, 300, 0, 0);
    // text corruptions of the .msh: digits turned into signs and letters
    const std::vector<unsigned char> msh = slurp(argv[4]);
    for (int k = 0; k < 300; k++) {
        std::vector<unsigned char> b = msh;
        const char subst[] = {'-', '9', 'x', ' ', '\n', '0'};
        b[next() % b.size()] = (unsigned char)subst[next() % sizeof subst];
        spill(dir + "/fuzz_msh_text", b);
        try_msh(dir + "/fuzz_msh_text");
    }
    std::printf("readers: %d accepted, %d refused\n", g_ok, g_err);
    return 0;
}
