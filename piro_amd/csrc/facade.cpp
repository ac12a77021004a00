// facade.cpp -- the reference's solver API and case driver on top of the C
// ABI (host code; every field operation is a device call through
// include/piro_hip.h).  The .ini reader and the Paraview writer it uses are
// in host_io.cpp.
#include "../../include/piro_solver.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <numeric>
#include <sstream>

#include "piro_internal.hpp"

namespace piro {

namespace {

void ck(int rc)
{
    if (rc != PIRO_OK) throw SolverError(std::string("piro: ") + piro_last_error());
}

// string_utilities::splitString (src/stringutilities.cpp:9-20)
std::vector<std::string> split(const std::string &s, char d)
{
    std::vector<std::string> out;
    size_t start = 0, end;
    while ((end = s.find(d, start)) != std::string::npos) {
        out.push_back(s.substr(start, end - start));
        start = end + 1;
    }
    out.push_back(s.substr(start));
    return out;
}

template <typename T>
std::vector<T> to_vec(const std::string &s)
{
    std::vector<T> v;
    for (auto &tok : split(s, ' ')) {
        std::istringstream iss(tok);
        T x;
        if (iss >> x) v.push_back(x);
    }
    return v;
}

int count_spaces(const std::string &s) { return (int)std::count(s.begin(), s.end(), ' '); }

int count_words(const std::string &s)
{
    std::istringstream iss(s);
    return (int)std::distance(std::istream_iterator<std::string>(iss), std::istream_iterator<std::string>());
}

int stoi_req(const IniReader &r, const char *sec, const char *key)
{
    std::string v = r.get(sec, key);
    try {
        return std::stoi(v);
    } catch (...) {
        throw SolverError(std::string("missing or invalid [") + sec + "] " + key);
    }
}

float stof_req(const IniReader &r, const char *sec, const char *key)
{
    std::string v = r.get(sec, key);
    try {
        return std::stof(v);
    } catch (...) {
        throw SolverError(std::string("missing or invalid [") + sec + "] " + key);
    }
}

// face lists of bc::initbc (src/bc.cpp:115-146)
std::vector<std::vector<int>> face_lists(const int *n)
{
    std::vector<std::vector<int>> idx(6);
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    for (int64_t ind = 0; ind < N; ++ind) {
        int kd = (int)(ind / ((int64_t)n[1] * n[0]));
        int jd = (int)((ind % ((int64_t)n[1] * n[0])) / n[0]);
        int id = (int)(ind % n[0]);
        bool xi = id != 0 && id != n[0] - 1, yi = jd != 0 && jd != n[1] - 1, zi = kd != 0 && kd != n[2] - 1;
        if (jd == 0 && xi && zi) idx[3].push_back((int)ind);
        else if (jd == n[1] - 1 && xi && zi) idx[2].push_back((int)ind);
        if (kd == 0 && xi && yi) idx[4].push_back((int)ind);
        else if (kd == n[2] - 1 && xi && yi) idx[5].push_back((int)ind);
        if (id == 0 && yi && zi) idx[1].push_back((int)ind);
        else if (id == n[0] - 1 && yi && zi) idx[0].push_back((int)ind);
    }
    return idx;
}

}  // namespace

// ------------------------------------------------------------------ Case

void *Case::alloc(size_t bytes)
{
    void *p = nullptr;
    ck(piro_malloc(ctx_, bytes, &p));
    owned_.push_back(p);
    return p;
}

Case::Case(const std::string &assets_dir, const std::string &case_name, const CaseOptions &o) : opt(o)
{
    IniReader reader(assets_dir + "/" + case_name);
    if (!reader.ok()) throw SolverError("cannot open case file " + assets_dir + "/" + case_name);
    casename = reader.get("File", "casename");
    simindex = stoi_req(reader, "File", "simindex");
    time_scheme = stoi_req(reader, "Schemes", "Time");
    space_scheme = stoi_req(reader, "Schemes", "Space");
    solver_scheme = stoi_req(reader, "Schemes", "Solver");
    auto nn = to_vec<unsigned>(reader.get("Mesh", "n"));
    auto ll = to_vec<float>(reader.get("Mesh", "l"));
    if (nn.size() != 3 || ll.size() != 3) throw SolverError("[Mesh] n and l need 3 values");
    for (int d = 0; d < 3; ++d) {
        n[d] = (int)nn[d] + 2;  // src/preprocess.cpp:72-74
        l[d] = ll[d];
    }
    N = (int64_t)n[0] * n[1] * n[2];
    // [Mesh] o and s land in the AMR struct's int fields (src/preprocess.cpp:95-101)
    auto oo = to_vec<float>(reader.get("Mesh", "o", "0 0 0"));
    auto ss = to_vec<float>(reader.get("Mesh", "s", "1 1 1"));
    if (oo.size() != 3 || ss.size() != 3) throw SolverError("[Mesh] o and s need 3 values");
    for (int d = 0; d < 3; ++d) {
        grid.n[d] = n[d];
        grid.origin[d] = (int)oo[d];
        grid.spacing[d] = (int)ss[d];
    }
    if (solver_scheme == 40) {
        IniReader rs(assets_dir + "/solver.ini");
        jli_tol = stof_req(rs, "JLI", "tolerance");
        jli_urf = stof_req(rs, "JLI", "URF");
        jli_maxiter = stoi_req(rs, "JLI", "max_iter");
    }
    timestep = stof_req(reader, "Solve", "Timestep");
    totaltime = stof_req(reader, "Solve", "TotalTime");
    save = stoi_req(reader, "Solve", "Save");
    data_type = stoi_req(reader, "Solve", "Data");
    // DENSE keeps N x N floats on the device (the reference's format):
    // refuse cases whose matrix would not fit in 64 GiB
    if (data_type == 1 && (double)N * (double)N * 4.0 > 64.0 * (1ull << 30))
        throw SolverError("Data = 1 (DENSE) needs N^2 floats; this grid is too large for it");
    if (data_type < 0 || data_type > 3) throw SolverError("Invalid DataType");  // src/process.cpp:93-96
    if (data_type == 2) {
        IniReader rh(assets_dir + "/Data/hashtable.ini");
        ht_lf = stof_req(rh, "Table", "LoadFactor");
        ht_a = stof_req(rh, "Table", "a");
        ht_b = stof_req(rh, "Table", "b");
        ht_c = stof_req(rh, "Table", "c");
    }
    ck(piro_ctx_create(0, &ctx_));
    init_fields(reader, assets_dir);
    work = (float *)alloc((size_t)N * 3 * sizeof(float));
    work2 = (float *)alloc((size_t)N * 3 * sizeof(float));
    init_bc(assets_dir);
    if (!opt.out_dir.empty()) {
        std::filesystem::create_directories(opt.out_dir + "/" + casename + "/mesh/level");
    }
}

Case::~Case()
{
    if (ctx_) {
        for (void *p : owned_) (void)piro_free(ctx_, p);
        (void)piro_ctx_destroy(ctx_);
    }
}

int Case::match(const std::string &var) const
{
    // process::matchscalartovar (src/process.cpp:20-28): 0 when absent
    for (size_t v = 0; v < names.size(); ++v)
        if (names[v] == var) return (int)v;
    return 0;
}

void Case::init_fields(const IniReader &reader, const std::string &assets_dir)
{
    // src/preprocess.cpp:111-173
    std::string cl = reader.get("Simulation", "Constants"), sl = reader.get("Simulation", "Scalars");
    auto cnames = split(cl, ' ');
    auto cvals = to_vec<float>(reader.get("Simulation", "Values"));
    auto snames = split(sl, ' ');
    std::string vl = reader.get("Simulation", "Vectors", "");
    auto vnames = split(vl, ' ');
    int n_const = count_spaces(cl) + 1, n_scal = count_spaces(sl) + 1, n_vec = count_words(vl);
    auto icfiles = split(reader.get("IC", "filename"), ' ');
    IniReader ic(assets_dir + "/IC/distribution.ini");
    auto make_ic = [&](int index, bool vector) -> std::vector<float> {
        // Piro::initialcondition (src/ic.cpp:34-99)
        std::vector<float> v;
        if (index >= (int)icfiles.size()) throw SolverError("[IC] filename has too few entries");
        const std::string &kind = icfiles[index];
        if (kind == "Gaussian") {
            v.assign(N, 0.0f);
            float scale = stof_req(ic, "Gaussian", "Scalefactor");
            auto sigma = to_vec<float>(ic.get("Gaussian", "Sigma"));
            auto mean = to_vec<float>(ic.get("Gaussian", "Median"));
            if (sigma.size() != 3 || mean.size() != 3) throw SolverError("[Gaussian] Sigma/Median need 3 values");
            float *d = (float *)alloc((size_t)N * sizeof(float));
            ck(piro_ic_gaussian(ctx_, n[0], n[1], n[2], l, mean.data(), sigma.data(), scale, d));
            ck(piro_memcpy_d2h(ctx_, v.data(), d, (size_t)N * sizeof(float)));
        } else if (kind == "UniformVector") {
            v.assign(N * 3, 0.0f);
            auto dir = to_vec<float>(ic.get("Vector", "Direction"));
            auto val = to_vec<float>(ic.get("Vector", "Value"));
            if (dir.size() != 3 || val.empty()) throw SolverError("[Vector] Direction/Value missing");
            for (int64_t i = 0; i < N; ++i) {
                v[i] = dir[0] * val[0];
                v[N + i] = dir[1] * val[0];
                v[2 * N + i] = dir[2] * val[0];
            }
        } else if (kind == "Coulomb") {
            v.assign(N, 0.0f);
            auto center = to_vec<float>(ic.get("Coulomb", "center"));
            float Z = stof_req(ic, "Coulomb", "Z");
            double e = stof_req(ic, "Coulomb", "e"), eps0 = stof_req(ic, "Coulomb", "epsilon_0");
            for (int x = 0; x < n[0]; ++x)
                for (int y = 0; y < n[1]; ++y)
                    for (int z = 0; z < n[2]; ++z) {
                        float c0 = x * l[0] / n[0], c1 = y * l[1] / n[1], c2 = z * l[2] / n[2];
                        float r = (float)std::pow(std::pow(c0 - center[0], 2) + std::pow(c1 - center[1], 2) +
                                                      std::pow(c2 - center[2], 2),
                                                  0.5);
                        v[x + (int64_t)y * n[0] + (int64_t)z * n[0] * n[1]] =
                            (float)(-(Z * 1E15 * std::pow(e, 2)) / (4 * M_PI * eps0 * (r + 1E-10)));
                    }
        } else {
            v.assign(vector ? 3 * N : N, 0.0f);  // unknown kinds leave zeros
        }
        return v;
    };
    for (int i = 0; i < n_const; ++i) {
        names.push_back(i < (int)cnames.size() ? cnames[i] : "");
        types.push_back(0);
        float *d = (float *)alloc((size_t)N * sizeof(float));
        std::vector<float> h(N, i < (int)cvals.size() ? cvals[i] : 0.0f);
        ck(piro_memcpy_h2d(ctx_, d, h.data(), (size_t)N * sizeof(float)));
        fields.push_back(d);
        host_values.push_back(std::move(h));
    }
    for (int i = 0; i < n_scal; ++i) {
        names.push_back(i < (int)snames.size() ? snames[i] : "");
        types.push_back(1);
        auto h = make_ic(i, false);
        float *d = (float *)alloc((size_t)N * 3 * sizeof(float));
        ck(piro_memcpy_h2d(ctx_, d, h.data(), (size_t)N * sizeof(float)));  // first N values
        fields.push_back(d);
        host_values.push_back(std::move(h));
    }
    for (int i = 0; i < n_vec; ++i) {
        names.push_back(i < (int)vnames.size() ? vnames[i] : "");
        types.push_back(2);
        auto h = make_ic(n_scal, true);
        if ((int64_t)h.size() < 3 * N) h.resize(3 * N, 0.0f);
        float *d = (float *)alloc((size_t)N * 3 * sizeof(float));
        ck(piro_memcpy_h2d(ctx_, d, h.data(), (size_t)N * 3 * sizeof(float)));
        fields.push_back(d);
        host_values.push_back(std::move(h));
    }
    if (fields.size() < 2) throw SolverError("the case needs at least one constant and one scalar");
}

void Case::init_bc(const std::string &assets_dir)
{
    // bc::readbc / initbc / setbc / opencl_initBC (src/bc.cpp:107-178, 67-78,
    // 20-36): BC values come from assets/setup.ini; the snapshot is field
    // slot 1 with the faces whose property resolves to slot 1 overwritten
    IniReader r(assets_dir + "/setup.ini");
    std::string type = r.get("BC", "type");
    if (count_spaces(type) <= 1) return;
    auto props = split(r.get("BC", "property"), ' ');
    auto vals = to_vec<float>(r.get("BC", "values"));
    if (props.size() < 6 || vals.size() < 6) throw SolverError("setup.ini [BC] needs 6 properties and values");
    // bc::setbc writes each face's value into the host values of the slot
    // its property names (slot 0 when the name is unknown); the snapshot is
    // slot 1's host values afterwards (bc::opencl_initBC)
    auto lists = face_lists(n);
    for (int f = 0; f < 6; ++f) {
        std::vector<float> &hv = host_values[match(props[f])];
        for (int idx : lists[f])
            if (idx < (int64_t)hv.size()) hv[idx] = vals[f];
    }
    std::vector<float> snap(host_values[1].begin(), host_values[1].begin() + N);
    snapshot = (float *)alloc((size_t)N * sizeof(float));
    ck(piro_memcpy_h2d(ctx_, snapshot, snap.data(), (size_t)N * sizeof(float)));
    has_bc = true;
}

std::vector<float> Case::download(int slot) const
{
    size_t cnt = types[slot] == 2 ? (size_t)N * 3 : (size_t)N;
    std::vector<float> h(cnt);
    ck(piro_memcpy_d2h(ctx_, h.data(), fields[slot], cnt * sizeof(float)));
    return h;
}

int Case::solve()
{
    // src/solve.cpp:11-48, float time accumulation
    process solver(*this);
    float time = 0.0f;
    int steps = 0;
    while (time < totaltime && (opt.max_steps < 0 || steps < opt.max_steps)) {
        switch (simindex) {
        case 0: {
            scalarMatrix UEqn(solver.ddt_r("T") = solver.laplacian("T", "Alpha"));
            UEqn.Solve(time);
            break;
        }
        case 1: {
            scalarMatrix UEqn(solver.ddt_r("U") = solver.vector("v") & solver.gradient("U"));
            UEqn.Solve(time);
            break;
        }
        case 2: {
            scalarMatrix UEqn(solver.ddt_r("rho") = solver.div("rho", "u"));
            UEqn.Solve(time);
            break;
        }
        default:
            throw SolverError("Invalid sim index");
        }
        time += timestep;
        ++steps;
    }
    ck(piro_ctx_synchronize(ctx_));
    return steps;
}

// --------------------------------------------------------------- process

Field process::ddt_r(const std::string &var)
{
    int t = c_->match(var);
    return Field{c_, t, c_->fields[t]};
}

Operator process::laplacian(const std::string &var1, const std::string &var2)
{
    (void)var1;
    Case &c = *c_;
    if (c.lap_init) return c.lap;
    const float norm = piro_lap_norm(c.l[0], c.n[0] - 2);
    int64_t nnz = 0;
    ck(piro_lap_nnz(c.n[0], c.n[1], c.n[2], &nnz));
    Operator op;
    op.c = c_;
    op.format = c.data_type;
    if (c.data_type == 1) {
        // Piro::DENSE::laplacian (src/operatorlaplacian.cpp:65-102): N x N,
        // no diffusivity scaling
        op.dense = (float *)c.alloc((size_t)c.N * (size_t)c.N * sizeof(float));
        ck(piro_assemble_laplacian_dense(c.ctx(), c.n[0], c.n[1], c.n[2], c.space_scheme, norm, op.dense, c.N));
        c.lap = op;
        c.lap_init = true;
        return op;
    }
    op.rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
    op.col = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
    op.val = (float *)c.alloc((size_t)nnz * sizeof(float));
    op.nnz = nnz;
    if (c.data_type == 3) {
        int32_t *rows = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
        ck(piro_assemble_laplacian_coo(c.ctx(), c.n[0], c.n[1], c.n[2], c.space_scheme, norm, rows, op.col, op.val));
        ck(piro_coo_row_ptr(c.ctx(), (int)c.N, nnz, rows, op.rp));
    } else {
        ck(piro_assemble_laplacian_csr(c.ctx(), c.n[0], c.n[1], c.n[2], c.space_scheme, norm, op.rp, op.col,
                                       op.val));
        // csrscale (src/process.cpp:70-72) is issued on the CSR path only
        if (c.data_type == 0 && c.opt.alpha_row_scaled)
            ck(piro_row_scale(c.ctx(), (int)c.N, op.rp, op.val, c.fields[c.match(var2)]));
    }
    if (c.data_type == 2) {
        int ts = 0;
        ck(piro_ht_table_size((int)nnz, (int)c.N, c.ht_a, c.ht_b, c.ht_c, c.ht_lf, &ts));
        op.ht_size = ts;
        op.ht_keys = (int32_t *)c.alloc((size_t)ts * sizeof(int32_t));
        op.ht_vals = (float *)c.alloc((size_t)ts * sizeof(float));
        ck(piro_ht_build(c.ctx(), (int)c.N, op.rp, op.col, op.val, ts, op.ht_keys, op.ht_vals));
        int64_t hn = 0;
        ck(piro_ht_count(c.ctx(), (int)c.N, op.ht_keys, ts, &hn));
        // the stepping CSR is the table unpacked in ascending column order
        op.col = (int32_t *)c.alloc((size_t)std::max<int64_t>(hn, 1) * sizeof(int32_t));
        op.val = (float *)c.alloc((size_t)std::max<int64_t>(hn, 1) * sizeof(float));
        op.rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
        ck(piro_ht_to_csr(c.ctx(), (int)c.N, op.ht_keys, op.ht_vals, ts, op.rp, op.col, op.val));
        op.nnz = hn;
    }
    c.lap = op;
    c.lap_init = true;
    return op;
}

Operator process::gradient(const std::string &var)
{
    (void)var;
    Case &c = *c_;
    if (c.data_type != 0) throw SolverError("Invalid DataType (gradient is CSR only, src/process.cpp:137-140)");
    if (!c.grad_init) {
        const float norm = piro_grad_norm(c.l[0], c.n[0] - 2);
        Operator *g[3] = {&c.grad_x, &c.grad_y, &c.grad_z};
        for (int d = 0; d < 3; ++d) {
            int64_t nnz = 0;
            ck(piro_grad_nnz(c.n[0], c.n[1], c.n[2], d, &nnz));
            Operator &op = *g[d];
            op.c = c_;
            op.rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
            op.col = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
            op.val = (float *)c.alloc((size_t)nnz * sizeof(float));
            op.nnz = nnz;
            ck(piro_assemble_gradient_csr(c.ctx(), c.n[0], c.n[1], c.n[2], c.space_scheme, d, norm, op.rp, op.col,
                                          op.val));
        }
        c.grad_init = true;
    }
    Operator tag;
    tag.c = c_;
    tag.format = -1;  // marks "the three gradient matrices"
    return tag;
}

VectorRef process::vector(const std::string &var)
{
    if (c_->data_type != 0) throw SolverError("Invalid DataType (vector is CSR only)");
    return VectorRef{c_, c_->match(var)};
}

Operator operator&(const VectorRef &v, const Operator &grad)
{
    // kernelmethods::csrgeam (src/kernelcsrgeam.cpp:9-86), once (DOT_INIT)
    Case &c = *v.c;
    if (grad.format != -1) throw SolverError("operator& expects solver.gradient(...)");
    if (!c.dot_init) {
        Operator &D = c.rhs;
        D.c = v.c;
        D.format = 0;
        D.rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
        int64_t nnz = 0;
        ck(piro_geam3_rowptr(c.ctx(), (int)c.N, c.grad_x.rp, c.grad_x.col, c.grad_y.rp, c.grad_y.col, c.grad_z.rp,
                             c.grad_z.col, D.rp, &nnz));
        D.col = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
        D.val = (float *)c.alloc((size_t)nnz * sizeof(float));
        D.nnz = nnz;
        ck(piro_geam3_fill(c.ctx(), (int)c.N, c.grad_x.rp, c.grad_x.col, c.grad_x.val, c.grad_y.rp, c.grad_y.col,
                           c.grad_y.val, c.grad_z.rp, c.grad_z.col, c.grad_z.val, c.fields[v.slot], D.rp, D.col,
                           D.val));
        c.dot_init = true;
    }
    return c.rhs;
}

Operator process::div(const std::string &var1, const std::string &var2)
{
    (void)var1;  // PIRO does not use it (src/process.cpp:173-178)
    Case &c = *c_;
    if (c.data_type != 0) throw SolverError("Invalid DataType (div is CSR only, src/process.cpp:207-210)");
    if (c.div_init) return c.div_rhs;
    // Piro::CSR::div (src/operatordivergenceCSR.cpp): the gradient stencils
    // with norm 2*(l0/float(n0-2)); SpMSCSR scales component d's entries by
    // var2's component d at their row; csrgeam_2 merges the three as the
    // concatenation-merge without factors
    const int slot = c.match(var2);
    if (c.types[slot] != 2) throw SolverError("div needs a vector field as its second argument");
    const float norm = piro_grad_norm(c.l[0], c.n[0] - 2);
    Operator g[3];
    for (int d = 0; d < 3; ++d) {
        int64_t nnz = 0;
        ck(piro_grad_nnz(c.n[0], c.n[1], c.n[2], d, &nnz));
        g[d].c = c_;
        g[d].rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
        g[d].col = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
        g[d].val = (float *)c.alloc((size_t)nnz * sizeof(float));
        g[d].nnz = nnz;
        ck(piro_assemble_gradient_csr(c.ctx(), c.n[0], c.n[1], c.n[2], c.space_scheme, d, norm, g[d].rp, g[d].col,
                                      g[d].val));
        ck(piro_row_scale(c.ctx(), (int)c.N, g[d].rp, g[d].val, c.fields[slot] + (size_t)d * c.N));
    }
    std::vector<float> ones(3 * (size_t)c.N, 1.0f);
    float *d_ones = (float *)c.alloc(ones.size() * sizeof(float));
    ck(piro_memcpy_h2d(c.ctx(), d_ones, ones.data(), ones.size() * sizeof(float)));
    Operator &D = c.div_rhs;
    D.c = c_;
    D.format = 0;
    D.rp = (int32_t *)c.alloc((size_t)(c.N + 1) * sizeof(int32_t));
    int64_t nnz = 0;
    ck(piro_geam3_rowptr(c.ctx(), (int)c.N, g[0].rp, g[0].col, g[1].rp, g[1].col, g[2].rp, g[2].col, D.rp, &nnz));
    D.col = (int32_t *)c.alloc((size_t)nnz * sizeof(int32_t));
    D.val = (float *)c.alloc((size_t)nnz * sizeof(float));
    D.nnz = nnz;
    ck(piro_geam3_fill(c.ctx(), (int)c.N, g[0].rp, g[0].col, g[0].val, g[1].rp, g[1].col, g[1].val, g[2].rp, g[2].col,
                       g[2].val, d_ones, D.rp, D.col, D.val));
    c.div_init = true;
    return D;
}

// ------------------------------------------------------- time integration

Field Field::operator=(const Operator &op)
{
    // CLBuffer::operator= (src/operatoroverload.cpp:20-100)
    Case &cs = *c;
    piro_ctx *x = cs.ctx();
    float *out = cs.work;
    if (cs.time_scheme == 11) {
        if (op.format == 0)
            ck(piro_fe_step_csr(x, cs.n[0], cs.n[1], cs.n[2], op.rp, op.col, op.val, op.nnz, cs.timestep, d,
                                cs.has_bc ? cs.snapshot : nullptr, out));
        else if (op.format == 1)
            ck(piro_fe_step_dense(x, (int)cs.N, op.dense, cs.N, cs.timestep, d, out));
        else
            ck(piro_fe_step_allrows(x, (int)cs.N, op.rp, op.col, op.val, op.nnz, cs.timestep, d, out));
    } else if (cs.time_scheme == 12 && cs.solver_scheme == 40) {
        if (op.format != 0) throw SolverError("Wrong data type (JLI needs CSR)");
        int it = 0;
        float res = 0;
        ck(piro_jli_solve(x, cs.n[0], cs.n[1], cs.n[2], op.rp, op.col, op.val, op.nnz, cs.timestep, d, cs.jli_urf,
                          cs.jli_tol, cs.jli_maxiter, out, cs.work2, &it, &res));
    } else if (cs.time_scheme == 12 && cs.solver_scheme == 27) {
        // backward Euler (I - dt L) x = b by the mixed-precision dense LU; the
        // reference instead eliminates L once and returns a zero field
        if (op.format == 1) throw SolverError("Wrong data type (the LU path takes CSR, HT or COO)");
        const int n = (int)cs.N;
        if (n > 16384) throw SolverError("the dense LU path is limited to N <= 16384");
        if (!cs.lu_init) {
            // A = I - dt L as a concatenation (duplicates summed on densify)
            std::vector<int32_t> irp(n + 1), icol(n);
            for (int i = 0; i <= n; ++i) irp[i] = i;
            for (int i = 0; i < n; ++i) icol[i] = i;
            std::vector<int32_t> erp(n + 1, 0);
            std::vector<float> ival(n, 1.0f), v3(3 * (size_t)n, 0.0f);
            for (int i = 0; i < n; ++i) { v3[i] = -cs.timestep; v3[n + i] = 1.0f; }
            auto up = [&](const void *h, size_t bytes) {
                void *dptr = cs.alloc(bytes);
                ck(piro_memcpy_h2d(x, dptr, h, bytes));
                return dptr;
            };
            int32_t *d_irp = (int32_t *)up(irp.data(), irp.size() * 4), *d_icol = (int32_t *)up(icol.data(), icol.size() * 4);
            float *d_ival = (float *)up(ival.data(), ival.size() * 4), *d_v3 = (float *)up(v3.data(), v3.size() * 4);
            int32_t *d_erp = (int32_t *)up(erp.data(), erp.size() * 4);
            int32_t *rp = (int32_t *)cs.alloc((size_t)(n + 1) * 4);
            int64_t nnz = 0;
            ck(piro_geam3_rowptr(x, n, op.rp, op.col, d_irp, d_icol, d_erp, d_icol, rp, &nnz));
            int32_t *col = (int32_t *)cs.alloc((size_t)nnz * 4);
            float *val = (float *)cs.alloc((size_t)nnz * 4);
            ck(piro_geam3_fill(x, n, op.rp, op.col, op.val, d_irp, d_icol, d_ival, d_erp, d_icol, d_ival, d_v3, rp, col,
                               val));
            cs.lu = (float *)cs.alloc((size_t)n * n * 4);
            ck(piro_csr_to_dense(x, n, n, rp, col, val, cs.lu, n));
            cs.lu_piv = (int32_t *)cs.alloc((size_t)n * 4);
            int info = 0;
            ck(piro_lu_factor_f32(x, n, cs.lu, n, cs.lu_piv, &info));
            cs.lu_init = true;
        }
        ck(piro_memcpy_d2d(x, out, d, (size_t)n * 4));
        ck(piro_lu_solve_f32(x, n, cs.lu, n, cs.lu_piv, 1, out, 1));
    } else {
        throw SolverError("unsupported [Schemes] Time/Solver combination");
    }
    return Field{c, slot, out};
}

void scalarMatrix::Solve(float currenttime)
{
    // src/process.cpp:221-249: commit into field slot 1, rewrite the faces,
    // export every Save steps
    Case &c = *f_.c;
    c.ts = (int)(currenttime / c.timestep);
    ck(piro_memcpy_d2d(c.ctx(), c.fields[1], f_.d, (size_t)c.N * sizeof(float)));
    if (c.has_bc) ck(piro_set_faces(c.ctx(), c.n[0], c.n[1], c.n[2], c.snapshot, c.fields[1]));
    if ((c.ts + 1) % c.save == 0 && !c.opt.out_dir.empty()) c.export_paraview(c.ts);
}

void Case::export_paraview(int step)
{
    // scalarMatrix::Solve reads slot 1 back before export_paraview
    // (src/process.cpp:238-247); the other slots print their host values
    const std::vector<float> slot1 = download(1);
    std::vector<VtkField> vf;
    for (size_t k = 0; k < fields.size(); ++k) {
        if (types[k] == 0) continue;
        vf.push_back(VtkField{names[k], types[k], k == 1 ? slot1.data() : host_values[k].data()});
    }
    write_paraview(opt.out_dir, casename, step, grid, vf);
}

}  // namespace piro

// -------------------------------------------------------------- C ABI

extern "C" int piro_case_run(const char *assets_dir, const char *case_name, const char *out_dir, int max_steps,
                             int alpha_row_scaled, int *steps_done, float *h_field, int64_t field_len)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(assets_dir && case_name, "assets_dir and case_name are required");
    piro::CaseOptions opt;
    opt.out_dir = out_dir ? out_dir : "";
    opt.max_steps = max_steps;
    opt.alpha_row_scaled = alpha_row_scaled != 0;
    try {
        piro::Case c(assets_dir, case_name, opt);
        int s = c.solve();
        if (steps_done) *steps_done = s;
        if (h_field) {
            auto f = c.download(1);
            PIRO_REQUIRE(field_len >= c.N, "field buffer too small");
            std::copy(f.begin(), f.begin() + c.N, h_field);
        }
    } catch (const piro::SolverError &e) {
        throw piro::Error(PIRO_EINVAL, e.what());
    }
    PIRO_API_END
}

extern "C" int piro_paraview_write(const char *out_dir, const char *case_name, int step, const int *h_dims3,
                                   const int *h_origin3, const int *h_spacing3, int n_fields,
                                   const char *const *names, const int *types, const float *const *h_fields)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(out_dir && case_name && h_dims3 && n_fields >= 0, "null argument");
    PIRO_REQUIRE(n_fields == 0 || (names && types && h_fields), "null field arrays");
    PIRO_REQUIRE(step >= 0, "step < 0");
    piro::VtkGrid g;
    for (int d = 0; d < 3; ++d) {
        PIRO_REQUIRE(h_dims3[d] >= 3, "padded dims must be >= 3");
        g.n[d] = h_dims3[d];
        g.origin[d] = h_origin3 ? h_origin3[d] : 0;
        g.spacing[d] = h_spacing3 ? h_spacing3[d] : 1;
    }
    std::vector<piro::VtkField> vf;
    for (int k = 0; k < n_fields; ++k) {
        PIRO_REQUIRE(types[k] >= 0 && types[k] <= 2, "field type must be 0, 1 or 2");
        PIRO_REQUIRE(types[k] == 0 || (names[k] && h_fields[k]), "null field");
        if (types[k] != 0) vf.push_back(piro::VtkField{names[k], types[k], h_fields[k]});
    }
    piro::write_paraview(out_dir, case_name, step, g, vf);
    PIRO_API_END
}
