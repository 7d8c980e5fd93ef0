// CaseFile.hpp -- a case described in an INI file, the form of the reference's
// legacy run configuration (interface/LEGACY/Run.py:7-47, 113-135: [consts]
// gamma, [ICs] r/u/v/p as expressions of the cell centre x, y, one [BC-<name>]
// section per physical name, [time-integrator] scheme/T/dt), read by hcfd_run.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace casefile {

// An arithmetic expression of x and y (numbers, x, y, pi, + - * / ^, unary
// minus, parentheses, comparisons < <= > >= == != giving 1 or 0, and the
// functions sqrt exp log sin cos tan abs min max pow); parsed once.
class Expr {
   public:
    explicit Expr(const std::string& text);
    double operator()(double x, double y) const;
    const std::string& text() const { return src; }

    struct Node;

   private:
    std::string src;
    std::shared_ptr<const Node> root;
};

struct BC {
    std::string name;   // physical name of the boundary colour
    int type = 0;       // hcfd_bctype
    std::vector<Expr> q;  // primitive BC state (rho, u, v, p), evaluated at x = y = 0; empty: the ICs at (0, 0)
};

struct Case {
    std::string mesh;           // .hmsh path (relative paths are taken from the case file's folder)
    double gamma = 1.4;
    int precision = 0;          // HCFD_FP64 / HCFD_MIXED
    bool hilbert = true;        // interior cell order: Hilbert (true) or reverse Cuthill-McKee
    std::vector<Expr> ic;       // r, u, v, p
    std::map<std::string, BC> bcs;
    int stages = 1;             // euler 1, ssprk2 2, ssprk3 3
    double T = 0.0, dt = 0.0, cfl = 0.0, dt_out = 0.0;
    std::string stem;           // output files <stem>_<k>.hsln
};

// Reads and checks a case file; throws std::runtime_error naming the file,
// section and key of the first problem.
Case read_case(const std::string& path);

// BC type names of the case file (symmetry, wall, sup-inlet, sub-inlet,
// sup-outlet, sub-outlet) -> hcfd_bctype.
int bc_type_of(const std::string& name);

}  // namespace casefile
