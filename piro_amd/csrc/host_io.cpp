// host_io.cpp -- the host-only input/output of the solver facade: the .ini
// reader and the Paraview writer.  Plain C++ with no HIP dependency, so the
// same file also builds under AddressSanitizer/UBSan for the host checks
// (`make -C piro_amd/csrc sanitize`).
#include "../../include/piro_solver.hpp"

#include <algorithm>
#include <charconv>
#include <exception>
#include <filesystem>
#include <fstream>
#include <system_error>
#include <thread>

namespace piro {

namespace {

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(' ');
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(' ');
    return s.substr(a, b - a + 1);
}

// one value as the reference's stream inserts a float (default format,
// precision 6): std::to_chars in general format with a precision is
// specified to print as printf("%.6g") in the "C" locale, which is what the
// stream does; returns the length written to buf (>= 16 bytes)
int fmt_float(char *buf, int cap, float v)
{
    const std::to_chars_result r = std::to_chars(buf, buf + cap, (double)v, std::chars_format::general, 6);
    return (int)(r.ptr - buf);
}

// the interior cells of one field x-fastest, space-separated.  z-slabs are
// formatted on up to 16 threads and joined in order (the text does not
// depend on the split).  A thread that cannot be started leaves its slabs to
// the calling thread; every started thread is joined, and an exception from
// any slab is rethrown after the joins.
std::string interior_text(const int *n, const float *v)
{
    const int nz = n[2] - 2;
    if (nz <= 0 || n[1] < 3 || n[0] < 3) return std::string();
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({nz, 16, (int)hc}));
    std::vector<std::string> part(nt);
    std::vector<std::exception_ptr> err(nt);
    auto slab = [&](int k) {
        try {
            const int z0 = 1 + (int)((int64_t)nz * k / nt), z1 = 1 + (int)((int64_t)nz * (k + 1) / nt);
            std::string &out = part[k];
            out.reserve((size_t)(z1 - z0) * (n[1] - 2) * (n[0] - 2) * 12);
            char buf[32];
            for (int z = z0; z < z1; ++z)
                for (int y = 1; y < n[1] - 1; ++y)
                    for (int x = 1; x < n[0] - 1; ++x) {
                        if (!out.empty()) out += ' ';
                        out.append(buf, fmt_float(buf, sizeof buf, v[x + (int64_t)y * n[0] + (int64_t)z * n[0] * n[1]]));
                    }
        } catch (...) {
            err[k] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    int started = 1;  // slab 0 is the calling thread's
    try {
        pool.reserve(nt - 1);
        for (; started < nt; ++started) pool.emplace_back(slab, started);
    } catch (...) {
        // no more threads: the remaining slabs run here
    }
    for (int k = started; k < nt; ++k) slab(k);
    slab(0);
    for (std::thread &th : pool) th.join();
    for (const std::exception_ptr &e : err)
        if (e) std::rethrow_exception(e);
    std::string out;
    size_t len = 0;
    for (const std::string &p : part) len += p.size() + 1;
    out.reserve(len);
    for (const std::string &p : part) {
        if (p.empty()) continue;
        if (!out.empty()) out += ' ';
        out += p;
    }
    return out;
}

std::string step_tag(int step)
{
    // 5-digit zero padding (src/paraview.cpp:108-109)
    std::string ts = std::to_string(step);
    return std::string(ts.size() < 5 ? 5 - ts.size() : 0, '0') + ts;
}

}  // namespace

// ------------------------------------------------------------ IniReader

IniReader::IniReader(const std::string &path)
{
    std::ifstream f(path);
    if (!f.is_open()) return;
    ok_ = true;
    std::string line, section;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line[0] == '[' && line.back() == ']') {
            section = line.substr(1, line.size() - 2);
        } else {
            size_t p = line.find('=');
            if (p != std::string::npos) data_[section][trim(line.substr(0, p))] = trim(line.substr(p + 1));
        }
    }
}

std::string IniReader::get(const std::string &section, const std::string &key, const std::string &def) const
{
    auto s = data_.find(section);
    if (s == data_.end()) return def;
    auto k = s->second.find(key);
    return k == s->second.end() ? def : k->second;
}

// ------------------------------------------------------------- Paraview

std::string writevti(const VtkGrid &g, const std::vector<VtkField> &fields)
{
    // src/paraview.cpp:12-103: extents of the interior, int origin/spacing,
    // boundary cells erased (bc::prepbc's list), "Float64" ascii arrays
    const int *n = g.n;
    const std::string ext = "0 " + std::to_string(n[0] - 2) + " 0 " + std::to_string(n[1] - 2) + " 0 " +
                            std::to_string(n[2] - 2);
    auto three = [](const int *v) {
        return std::to_string(v[0]) + " " + std::to_string(v[1]) + " " + std::to_string(v[2]);
    };
    std::string s = "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\">\n";
    s += "<ImageData WholeExtent=\"" + ext + "\" Origin=\"" + three(g.origin) + "\" Spacing=\"" + three(g.spacing) +
         "\">\n";
    s += "<Piece Extent=\"" + ext + "\">\n";
    s += "<PointData Scalars=\"ScalarValues\">\n</PointData>\n<CellData>\n";
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    for (const VtkField &f : fields) {
        if (f.type == 1) {
            s += "<DataArray type=\"Float64\" Name=\"" + f.name + "\" format=\"ascii\">\n";
            s += interior_text(n, f.data) + "</DataArray>\n";
        } else if (f.type == 2) {
            const char *suf[3] = {"_x", "_y", "_z"};
            for (int d = 0; d < 3; ++d) {
                s += "<DataArray type=\"Float64\" Name=\"" + f.name + suf[d] + "\" format=\"ascii\">\n";
                s += interior_text(n, f.data + d * N) + "</DataArray>\n";
            }
        }
    }
    s += "</CellData>\n</Piece>\n</ImageData>\n</VTKFile>\n";
    return s;
}

std::string writevth(const VtkGrid &g, int step)
{
    // src/paraview.cpp:105-137 (one level, one block)
    const std::string ts = step_tag(step);
    std::string s = "<?xml version=\"1.0\"?>\n";
    s += "<VTKFile type=\"vtkOverlappingAMR\" version=\"1.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n";
    s += "<vtkOverlappingAMR origin=\"0 0 0\" grid_description=\"XYZ\">\n";
    s += "<Block level=\"0\" spacing=\"1 1 1\">\n";
    s += "<DataSet index=\"0\" amr_box=\"0 " + std::to_string(g.n[0] - 2) + " 0 " + std::to_string(g.n[1] - 2) +
         " 0 " + std::to_string(g.n[2] - 2) + "\" file=\"level/" + ts + "_level_0.vti\">\n";
    s += "</DataSet>\n</Block>\n</vtkOverlappingAMR>\n</VTKFile>\n";
    return s;
}

void write_paraview(const std::string &out_dir, const std::string &casename, int step, const VtkGrid &g,
                    const std::vector<VtkField> &fields)
{
    const std::string base = out_dir + "/" + casename + "/mesh/";
    std::filesystem::create_directories(base + "level");
    const std::string ts = step_tag(step);
    std::ofstream vti(base + "level/" + ts + "_level_0.vti");
    vti << writevti(g, fields) << "\n";
    std::ofstream vth(base + "mesh_" + ts + ".vth");
    vth << writevth(g, step) << "\n";
    if (!vti || !vth) throw SolverError("cannot write Paraview files under " + base);
}

}  // namespace piro
