// piro_solver.hpp -- C++ drop-in for PIRO's solver API on the MI355X path.
//
// Mirrors the reference's operator layer so that its time loop
// (src/solve.cpp:18-46) reads the same:
//
//     piro::process solver(case_);
//     piro::scalarMatrix UEqn(solver.ddt_r("T") = solver.laplacian("T", "Alpha"));
//     UEqn.Solve(time);
//
//     piro::scalarMatrix VEqn(solver.ddt_r("U") = solver.vector("v") & solver.gradient("U"));
//
// Reference interfaces replaced (file:line, PIRO tree):
//   Piro::preprocess                 src/preprocess.cpp:20-208
//   Piro::solve                      src/solve.cpp:11-48
//   process::ddt_r/laplacian/gradient/vector   src/process.cpp:52-166,
//                                    include/process.hpp:47-81
//   CLBuffer::operator= / operator&  src/operatoroverload.cpp:20-123,
//                                    include/operatoroverload.hpp:15,19
//   scalarMatrix::Solve              src/process.cpp:217-249
//   paraview::writevth/writevti      src/paraview.cpp:12-137
//
// Device work goes through the C ABI of include/piro_hip.h.  Errors are
// reported as piro::SolverError exceptions inside C++ and as status codes
// through piro_case_run (piro_hip.h).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "piro_hip.h"

namespace piro {

struct SolverError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// [section] key = value reader with the reference's parsing rules
// (src/fileutilities.cpp:124-175): trims spaces, skips ';'/'#' comments.
class IniReader {
  public:
    explicit IniReader(const std::string &path);
    std::string get(const std::string &section, const std::string &key,
                    const std::string &def = "default_value") const;
    bool ok() const { return ok_; }

  private:
    std::map<std::string, std::map<std::string, std::string>> data_;
    bool ok_ = false;
};

// ------------------------------------------------------------ Paraview
// src/paraview.cpp:12-137 with one level and one block ([Mesh] levels = 1,
// index = 1; the reference's writer indexes past its cell data otherwise).

struct VtkGrid {
    int n[3] = {0, 0, 0};        // padded dims
    int origin[3] = {0, 0, 0};   // [Mesh] o, truncated to int as the reference's AMR struct stores it
    int spacing[3] = {1, 1, 1};  // [Mesh] s, likewise
};
struct VtkField {
    std::string name;
    int type = 1;                // 0 constant (never written), 1 scalar (N values), 2 vector (3N values)
    const float *data = nullptr; // host
};
// the .vti text: interior cells only, each scalar and each vector component
// as one ascii "Float64" DataArray of the values printed as std::ostream
// prints a float
std::string writevti(const VtkGrid &g, const std::vector<VtkField> &fields);
// the .vth index pointing at level/<step>_level_0.vti
std::string writevth(const VtkGrid &g, int step);
// writes <out_dir>/<casename>/mesh/mesh_<step>.vth and
// <out_dir>/<casename>/mesh/level/<step>_level_0.vti (each followed by a
// newline, src/fileutilities.cpp:24-40), creating the folders
void write_paraview(const std::string &out_dir, const std::string &casename, int step, const VtkGrid &g,
                    const std::vector<VtkField> &fields);

class Case;

// A device field slot (the reference's CLBuffer with its `ind`).
struct Field {
    Case *c = nullptr;
    int slot = -1;
    float *d = nullptr;   // device data (N or 3N floats)
    // CLBuffer::operator=: advance one time step of `op` from this field and
    // return the new state (not yet committed; scalarMatrix::Solve commits)
    Field operator=(const struct Operator &op);
};

// An assembled operator (CSR on the device; the HT/COO formats are kept in
// their own layout and unpacked to the CSR the step kernels read).
struct Operator {
    Case *c = nullptr;
    int format = 0;              // 0 CSR, 1 DENSE, 2 HT, 3 COO
    int32_t *rp = nullptr, *col = nullptr;
    float *val = nullptr;
    int64_t nnz = 0;
    int32_t *ht_keys = nullptr;  // HT only
    float *ht_vals = nullptr;
    int ht_size = 0;
    float *dense = nullptr;      // DENSE only: N x N row-major
};

struct VectorRef {
    Case *c = nullptr;
    int slot = -1;
};

Operator operator&(const VectorRef &v, const Operator &grad);

class process {
  public:
    explicit process(Case &c) : c_(&c) {}
    Field ddt_r(const std::string &var);
    Operator laplacian(const std::string &var1, const std::string &var2);
    Operator gradient(const std::string &var);
    VectorRef vector(const std::string &var);
    // process::div (src/process.cpp:168-212), CSR only: the divergence
    // matrices with each component's rows scaled by var2's component
    Operator div(const std::string &var1, const std::string &var2);

  private:
    Case *c_;
};

class scalarMatrix {
  public:
    explicit scalarMatrix(Field f) : f_(f) {}
    void Solve(float currenttime);

  private:
    Field f_;
};

struct CaseOptions {
    std::string out_dir;        // where <casename>/mesh/... is written ("" = none)
    bool alpha_row_scaled = true;  // false: reproduce the shipped (alpha-free) operator
    int max_steps = -1;         // cap on time steps (-1: until TotalTime)
};

// The case: parsed .ini files, device fields, lazily assembled operators.
class Case {
  public:
    Case(const std::string &assets_dir, const std::string &case_name, const CaseOptions &opt);
    ~Case();
    Case(const Case &) = delete;
    Case &operator=(const Case &) = delete;

    // time loop of src/solve.cpp:18-46; returns the number of steps taken
    int solve();

    // accessors used by the operator layer
    piro_ctx *ctx() const { return ctx_; }
    int n[3] = {0, 0, 0};       // padded dims
    int64_t N = 0;
    float l[3] = {0, 0, 0};
    VtkGrid grid;               // dims, [Mesh] o and s for the Paraview files
    float timestep = 0, totaltime = 0;
    int save = 1, simindex = 0, time_scheme = 11, space_scheme = 3, solver_scheme = 0, data_type = 0;
    std::string casename;
    std::vector<std::string> names;   // slot names (constants, scalars, vectors)
    std::vector<int> types;           // 0 constant, 1 scalar, 2 vector
    std::vector<float *> fields;      // device data per slot
    // host values per slot as the reference's CellData keeps them: the
    // initial condition with the BC face values written into the slot each
    // BC property names; the Paraview files print slot 1 from the device
    // and every other slot from here
    std::vector<std::vector<float>> host_values;
    float *snapshot = nullptr;        // BC snapshot (field slot 1 with face values)
    float *work = nullptr;            // step output buffer
    float *work2 = nullptr;
    bool has_bc = false;
    float jli_tol = 1e-5f, jli_urf = 1.0f;
    int jli_maxiter = 1000;
    float ht_a = 0, ht_b = 0, ht_c = 0, ht_lf = 0.7f;
    int ts = 0;
    CaseOptions opt;
    int match(const std::string &var) const;  // matchscalartovar: 0 if absent
    Operator lap, grad_x, grad_y, grad_z, rhs, div_rhs;
    bool lap_init = false, grad_init = false, dot_init = false, lu_init = false, div_init = false;
    float *lu = nullptr, *lu_a = nullptr;
    int32_t *lu_piv = nullptr;
    void export_paraview(int step);
    std::vector<float> download(int slot) const;
    void *alloc(size_t bytes);  // device memory owned by the case

  private:
    piro_ctx *ctx_ = nullptr;
    std::vector<void *> owned_;
    void init_fields(const IniReader &reader, const std::string &assets_dir);
    void init_bc(const std::string &assets_dir);
    friend struct Field;
};


}  // namespace piro
