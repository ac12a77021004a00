// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own host-side assembly code (compiled from the PIRO
// sources by oracle/Makefile into oracle/_ref/) on small grids and dumps the
// results as raw little-endian arrays, so tests/golden/make_fixtures.py can
// freeze them as fixtures.  Nothing here is part of the product.
//
// Reference entry points exercised:
//   Piro::CSR::stencils::laplacian_{cd,uw,dw}   src/stencils.cpp:233-429
//   Piro::CSR::stencils::gradient_{cd,uw,dw}_*  src/stencils.cpp:7-231
//   Piro::DENSE::stencils::laplacian_{cd,uw,dw} src/stencils.cpp:431-571
//   Piro::COO::stencils::laplacian_{cd,uw,dw}   src/stencils.cpp:572-795
//   Piro::HT::stencils::laplacian_{cd,uw,dw}    src/stencils.cpp:797-993
//     (+ the sethash loop of HT::laplacian,      src/operatorlaplacian.cpp:194-204)
//   Piro::methods::sethash                      src/methods.cpp:74-118
//   Piro::calculategaussian                     src/ic.cpp:16-27
//   Piro::calculatecoulomb                      src/ic.cpp:29-32 (coordinates
//     and std::stof parameter parsing as in     src/ic.cpp:62-79)
//   Piro::bc::prepbc + Piro::paraview::writevth src/bc.cpp:80-105,
//     (+ writevti, writefile)                    src/paraview.cpp:12-137,
//                                                src/fileutilities.cpp:24-40
//
// Usage:
//   piro_ref_harness lap  <scheme 1|2|3> nx ny nz norm_hex out.bin
//   piro_ref_harness grad <scheme> <dir 0|1|2> nx ny nz norm_hex out.bin
//   piro_ref_harness dense <scheme> nx ny nz norm_hex out.bin  (N x N floats)
//   piro_ref_harness ht   <scheme> nx ny nz norm_hex table_size out.bin
//   piro_ref_harness coo  <scheme> nx ny nz norm_hex out.bin     (rows, cols, vals)
//   piro_ref_harness htst <scheme> nx ny nz norm_hex table_size out.bin
//     (HT stencil row_ptr, columns, values, then the table keys, values)
//   piro_ref_harness gauss nx ny nz lx ly lz mx my mz sx sy sz out.bin
//   piro_ref_harness coulomb nx ny nz lx ly lz cx cy cz Z e eps0 out.bin
//     (Z, e, eps0 as text, parsed with std::stof like the reference)
//   piro_ref_harness vtk  nx ny nz n_scal n_vec step in.bin out_dir
//     in.bin: n_scal fields of nx*ny*nz floats, then n_vec fields of
//     3*nx*ny*nz floats (raw little-endian); fields are named S0.., V0..
//     after one constant "Alpha"; writes out_dir/case/mesh/mesh_<step>.vth
//     and out_dir/case/mesh/level/<step>_level_0.vti with the reference's
//     own writer
// norm_hex is the IEEE-754 bit pattern of the float norm (so no decimal
// round trip can perturb it).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <filesystem>

#include <bc.hpp>
#include <datatypes.hpp>
#include <fileutilities.hpp>
#include <ic.hpp>
#include <postprocess.hpp>
#include <methods.hpp>
#include <stencils.hpp>

static float from_hex(const char *s)
{
    uint32_t u = (uint32_t)strtoul(s, nullptr, 16);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template <typename T>
static void put(FILE *f, const std::vector<T> &v)
{
    int64_t n = (int64_t)v.size();
    fwrite(&n, sizeof(n), 1, f);
    if (n) fwrite(v.data(), sizeof(T), (size_t)n, f);
}

static void set_mesh(unsigned nx, unsigned ny, unsigned nz)
{
    Piro::MeshParams &MP = Piro::MeshParams::getInstance();
    MP.setvalue(Piro::MeshParams::num_cells, std::vector<uint>{nx, ny, nz});
}

static void lap(int scheme, Piro::CellData &cd, std::vector<uint> n, float norm)
{
    cd.rowpointers.assign(n[0] * n[1] * n[2] + 1, 0);
    if (scheme == 3) Piro::CSR::stencils::laplacian_cd(cd, n, norm);
    else if (scheme == 2) Piro::CSR::stencils::laplacian_dw(cd, n, norm);
    else Piro::CSR::stencils::laplacian_uw(cd, n, norm);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    std::string cmd = argv[1];
    if (cmd == "lap" && argc == 8) {
        int scheme = atoi(argv[2]);
        std::vector<uint> n = {(uint)atoi(argv[3]), (uint)atoi(argv[4]), (uint)atoi(argv[5])};
        set_mesh(n[0], n[1], n[2]);
        Piro::CellData cd;
        lap(scheme, cd, n, from_hex(argv[6]));
        FILE *f = fopen(argv[7], "wb");
        put(f, cd.rowpointers); put(f, cd.columns); put(f, cd.values);
        fclose(f);
        return 0;
    }
    if (cmd == "grad" && argc == 9) {
        int scheme = atoi(argv[2]), dir = atoi(argv[3]);
        std::vector<uint> n = {(uint)atoi(argv[4]), (uint)atoi(argv[5]), (uint)atoi(argv[6])};
        set_mesh(n[0], n[1], n[2]);
        float norm = from_hex(argv[7]);
        Piro::CellData cd;
        cd.rowpointers.assign(n[0] * n[1] * n[2] + 1, 0);
        using F = void (*)(Piro::CellData &, std::vector<uint>, float);
        F tab[3][3] = {
            {Piro::CSR::stencils::gradient_uw_x, Piro::CSR::stencils::gradient_uw_y, Piro::CSR::stencils::gradient_uw_z},
            {Piro::CSR::stencils::gradient_dw_x, Piro::CSR::stencils::gradient_dw_y, Piro::CSR::stencils::gradient_dw_z},
            {Piro::CSR::stencils::gradient_cd_x, Piro::CSR::stencils::gradient_cd_y, Piro::CSR::stencils::gradient_cd_z}};
        tab[scheme - 1][dir](cd, n, norm);
        FILE *f = fopen(argv[8], "wb");
        put(f, cd.rowpointers); put(f, cd.columns); put(f, cd.values);
        fclose(f);
        return 0;
    }
    if (cmd == "dense" && argc == 8) {
        int scheme = atoi(argv[2]);
        std::vector<uint> n = {(uint)atoi(argv[3]), (uint)atoi(argv[4]), (uint)atoi(argv[5])};
        set_mesh(n[0], n[1], n[2]);
        const size_t N = (size_t)n[0] * n[1] * n[2];
        Piro::CellData cd;
        cd.matrix.assign(N * N, 0.0f);
        const float norm = from_hex(argv[6]);
        if (scheme == 3) Piro::DENSE::stencils::laplacian_cd(cd, n, norm);
        else if (scheme == 2) Piro::DENSE::stencils::laplacian_dw(cd, n, norm);
        else Piro::DENSE::stencils::laplacian_uw(cd, n, norm);
        FILE *f = fopen(argv[7], "wb");
        put(f, cd.matrix);
        fclose(f);
        return 0;
    }
    if (cmd == "ht" && argc == 9) {
        int scheme = atoi(argv[2]);
        std::vector<uint> n = {(uint)atoi(argv[3]), (uint)atoi(argv[4]), (uint)atoi(argv[5])};
        set_mesh(n[0], n[1], n[2]);
        Piro::CellData cd;
        lap(scheme, cd, n, from_hex(argv[6]));
        int ts = atoi(argv[7]);
        int N = (int)(n[0] * n[1] * n[2]);
        std::vector<int> keys(ts, -1);
        std::vector<float> vals(ts, 0.0f);
        for (int i = 0; i < N; i++)
            for (int j = cd.rowpointers[i]; j < cd.rowpointers[i + 1]; j++)
                Piro::methods::sethash(i * N + cd.columns[j], cd.values[j], ts, keys, vals);
        FILE *f = fopen(argv[8], "wb");
        put(f, keys); put(f, vals);
        fclose(f);
        return 0;
    }
    if (cmd == "coo" && argc == 8) {
        int scheme = atoi(argv[2]);
        std::vector<uint> n = {(uint)atoi(argv[3]), (uint)atoi(argv[4]), (uint)atoi(argv[5])};
        set_mesh(n[0], n[1], n[2]);
        Piro::CellData cd;
        const float norm = from_hex(argv[6]);
        if (scheme == 3) Piro::COO::stencils::laplacian_cd(cd, n, norm);
        else if (scheme == 2) Piro::COO::stencils::laplacian_dw(cd, n, norm);
        else Piro::COO::stencils::laplacian_uw(cd, n, norm);
        FILE *f = fopen(argv[7], "wb");
        put(f, cd.rows); put(f, cd.columns); put(f, cd.values);
        fclose(f);
        return 0;
    }
    if (cmd == "htst" && argc == 9) {
        int scheme = atoi(argv[2]);
        std::vector<uint> n = {(uint)atoi(argv[3]), (uint)atoi(argv[4]), (uint)atoi(argv[5])};
        set_mesh(n[0], n[1], n[2]);
        const int N = (int)(n[0] * n[1] * n[2]);
        Piro::CellData cd;
        cd.rowpointers.assign(N + 1, 0);  // as HT::laplacian does before the stencil
        const float norm = from_hex(argv[6]);
        if (scheme == 3) Piro::HT::stencils::laplacian_cd(cd, n, norm);
        else if (scheme == 2) Piro::HT::stencils::laplacian_dw(cd, n, norm);
        else Piro::HT::stencils::laplacian_uw(cd, n, norm);
        const int ts = atoi(argv[7]);
        std::vector<int> keys(ts, -1);
        std::vector<float> vals(ts, 0.0f);
        for (int i = 0; i < N; i++)
            for (int j = cd.rowpointers[i]; j < cd.rowpointers[i + 1]; j++)
                Piro::methods::sethash(i * N + cd.columns[j], cd.values[j], ts, keys, vals);
        FILE *f = fopen(argv[8], "wb");
        put(f, cd.rowpointers); put(f, cd.columns); put(f, cd.values);
        put(f, keys); put(f, vals);
        fclose(f);
        return 0;
    }
    if (cmd == "coulomb" && argc == 15) {
        std::vector<uint> n = {(uint)atoi(argv[2]), (uint)atoi(argv[3]), (uint)atoi(argv[4])};
        std::vector<float> l = {(float)atof(argv[5]), (float)atof(argv[6]), (float)atof(argv[7])};
        std::vector<float> center = {(float)atof(argv[8]), (float)atof(argv[9]), (float)atof(argv[10])};
        float Z = std::stof(argv[11]);
        double e = std::stof(argv[12]);
        double eps0 = std::stof(argv[13]);
        std::vector<float> out((size_t)n[0] * n[1] * n[2]);
        std::vector<float> c(3);
        for (int x = 0; x < (int)n[0]; ++x)
            for (int y = 0; y < (int)n[1]; ++y)
                for (int z = 0; z < (int)n[2]; ++z) {
                    c[0] = x * l[0] / n[0];
                    c[1] = y * l[1] / n[1];
                    c[2] = z * l[2] / n[2];
                    out[x + (size_t)y * n[0] + (size_t)z * n[0] * n[1]] =
                        Piro::calculatecoulomb(c, center, Z, e, eps0);
                }
        FILE *f = fopen(argv[14], "wb");
        put(f, out);
        fclose(f);
        return 0;
    }
    if (cmd == "gauss" && argc == 15) {
        int n[3] = {atoi(argv[2]), atoi(argv[3]), atoi(argv[4])};
        float l[3] = {(float)atof(argv[5]), (float)atof(argv[6]), (float)atof(argv[7])};
        std::vector<float> mean = {(float)atof(argv[8]), (float)atof(argv[9]), (float)atof(argv[10])};
        std::vector<float> sigma = {(float)atof(argv[11]), (float)atof(argv[12]), (float)atof(argv[13])};
        std::vector<float> out((size_t)n[0] * n[1] * n[2]);
        std::vector<float> c(3);
        for (int x = 0; x < n[0]; ++x)
            for (int y = 0; y < n[1]; ++y)
                for (int z = 0; z < n[2]; ++z) {
                    c[0] = x * l[0] / (float)n[0];
                    c[1] = y * l[1] / (float)n[1];
                    c[2] = z * l[2] / (float)n[2];
                    out[x + (size_t)y * n[0] + (size_t)z * n[0] * n[1]] =
                        Piro::calculategaussian(c, mean, sigma);
                }
        FILE *f = fopen(argv[14], "wb");
        put(f, out);
        fclose(f);
        return 0;
    }
    if (cmd == "vtk" && argc == 10) {
        std::vector<uint> n = {(uint)atoi(argv[2]), (uint)atoi(argv[3]), (uint)atoi(argv[4])};
        const int n_scal = atoi(argv[5]), n_vec = atoi(argv[6]), step = atoi(argv[7]);
        const size_t N = (size_t)n[0] * n[1] * n[2];
        Piro::MeshParams &MP = Piro::MeshParams::getInstance();
        Piro::SolveParams &SP = Piro::SolveParams::getInstance();
        MP.setvalue(Piro::MeshParams::num_cells, n);
        MP.setvalue(Piro::MeshParams::VECTORNUM, n_vec);
        MP.setvalue(Piro::MeshParams::LEVELS, 1);
        MP.setvalue(Piro::MeshParams::INDEX, std::vector<int>{1});
        SP.setvalue(Piro::SolveParams::CASENAME, std::string("case"));
        Piro::AMR amr;
        for (int d = 0; d < 3; ++d) {
            amr.WholeExtent[2 * d] = 0;
            amr.WholeExtent[2 * d + 1] = (int)n[d];
            amr.Origin[d] = 0;
            amr.Spacing[d] = 1;
        }
        FILE *f = fopen(argv[8], "rb");
        if (!f) return 3;
        Piro::CellData cd;
        cd.Scalars = "Alpha";
        cd.type = 0;
        cd.values.assign(N, 0.1f);
        amr.CD.push_back(cd);
        for (int i = 0; i < n_scal + n_vec; ++i) {
            const bool vec = i >= n_scal;
            cd.Scalars = vec ? "V" + std::to_string(i - n_scal) : "S" + std::to_string(i);
            cd.type = vec ? 2 : 1;
            cd.values.assign(vec ? 3 * N : N, 0.0f);
            if (fread(cd.values.data(), sizeof(float), cd.values.size(), f) != cd.values.size()) return 3;
            amr.CD.push_back(cd);
        }
        fclose(f);
        // the three operator slots preprocess appends (src/preprocess.cpp:197-202)
        Piro::CellData op;
        for (int i = 0; i < 3; ++i) amr.CD.push_back(op);
        MP.setvalue(Piro::MeshParams::AMR, std::vector<Piro::AMR>{amr});
        Piro::bc::prepbc();
        Piro::file_utilities::current_path = argv[9];
        std::filesystem::create_directories(std::string(argv[9]) + "/case/mesh/level");
        Piro::paraview::writevth(step);
        return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 2;
}
