// host_io_check.cpp -- drives the host-only input/output of the library
// (host_io.cpp: the .ini reader and the Paraview writer) for the
// AddressSanitizer/UBSan build of piro_amd/csrc/sanitize/Makefile.
//
//   host_io_check ini <path> <section> <key> [<section> <key> ...]
//       prints one value per pair (the reader's default for a missing key)
//   host_io_check vtk nx ny nz n_scal n_vec step in.bin out_dir
//       in.bin: n_scal fields of nx*ny*nz floats, then n_vec fields of
//       3*nx*ny*nz floats; fields are named S0.., V0..; writes
//       out_dir/case/mesh/mesh_<step>.vth and level/<step>_level_0.vti
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "piro_solver.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "ini" && argc >= 3 && argc % 2 == 1) {
            piro::IniReader r(argv[2]);
            if (!r.ok()) return 3;
            for (int i = 3; i + 1 < argc; i += 2) std::printf("%s\n", r.get(argv[i], argv[i + 1]).c_str());
            return 0;
        }
        if (cmd == "vtk" && argc == 10) {
            piro::VtkGrid g;
            for (int d = 0; d < 3; ++d) g.n[d] = std::atoi(argv[2 + d]);
            const int n_scal = std::atoi(argv[5]), n_vec = std::atoi(argv[6]), step = std::atoi(argv[7]);
            const size_t N = (size_t)g.n[0] * g.n[1] * g.n[2];
            FILE *f = std::fopen(argv[8], "rb");
            if (!f) return 3;
            std::vector<std::vector<float>> data(n_scal + n_vec);
            std::vector<piro::VtkField> fields;
            for (int k = 0; k < n_scal + n_vec; ++k) {
                const bool vec = k >= n_scal;
                data[k].resize(vec ? 3 * N : N);
                if (std::fread(data[k].data(), sizeof(float), data[k].size(), f) != data[k].size()) return 3;
                fields.push_back(piro::VtkField{vec ? "V" + std::to_string(k - n_scal) : "S" + std::to_string(k),
                                                vec ? 2 : 1, data[k].data()});
            }
            std::fclose(f);
            piro::write_paraview(argv[9], "case", step, g, fields);
            return 0;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
