#pragma once
// AMReX plotfile writer (WriteMultiLevelPlotfile's on-disk format,
// "HyperCLaw-V1.1"): Header, then per level Level_<l>/Cell_H and one
// Cell_D_00000 data file holding every box as a FAB (one header line, then
// little-endian float64, component-major, i fastest).  Variables are var0.. then aux0..
// (Source/ZoomyAmr.cpp:851-857); this is the layout deliverable.py:33-87 reads.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace zamr {

struct PlotBox {
    int lo[2], hi[2];
    std::vector<double> data;  // (ncomp, ny_box, nx_box)
};

inline bool make_dir(const std::string& d) {
    if (mkdir(d.c_str(), 0755) == 0) return true;
    return errno == EEXIST;
}

// Concatenate(Concatenate("plt_", identifier), step, 5) (ZoomyAmr.cpp:824):
// Concatenate pads its number to at least 5 digits, so both numbers are
// padded (identifier 0, step 12 -> plt_0000000012).
inline std::string plotfile_name(int identifier, int step) {
    std::ostringstream s;
    s << "plt_" << std::setw(5) << std::setfill('0') << identifier << std::setw(5) << std::setfill('0') << step;
    return s.str();
}

struct PlotLevel {
    int nx, ny;        // level domain (0, 0) .. (nx-1, ny-1)
    double dx[2];
    std::vector<PlotBox> boxes;
};

// One level's Cell_D_00000 + Cell_H under path/Level_<lev>.
inline bool write_plot_level(const std::string& path, int lev, int ncomp, const std::vector<PlotBox>& boxes) {
    const std::string dir = path + "/Level_" + std::to_string(lev);
    if (!make_dir(dir)) return false;
    std::vector<long long> offsets;
    std::vector<double> mins, maxs;
    {
        std::ofstream d(dir + "/Cell_D_00000", std::ios::binary);
        if (!d) return false;
        for (auto& b : boxes) {
            offsets.push_back((long long)d.tellp());
            std::ostringstream fh;
            fh << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))((" << b.lo[0] << "," << b.lo[1]
               << ") (" << b.hi[0] << "," << b.hi[1] << ") (0,0)) " << ncomp << "\n";
            const std::string s = fh.str();
            d.write(s.data(), (std::streamsize)s.size());
            d.write(reinterpret_cast<const char*>(b.data.data()), (std::streamsize)(b.data.size() * sizeof(double)));
            const size_t per = b.data.size() / (size_t)ncomp;
            for (int c = 0; c < ncomp; ++c) {
                double mn = std::numeric_limits<double>::infinity(), mx = -mn;
                for (size_t k = 0; k < per; ++k) {
                    const double v = b.data[c * per + k];
                    mn = v < mn ? v : mn;
                    mx = v > mx ? v : mx;
                }
                mins.push_back(mn);
                maxs.push_back(mx);
            }
        }
        if (!d) return false;
    }
    std::ofstream c(dir + "/Cell_H");
    if (!c) return false;
    c << std::setprecision(17);
    c << 1 << "\n" << 1 << "\n" << ncomp << "\n" << 0 << "\n";
    c << "(" << boxes.size() << " 0\n";
    for (auto& b : boxes)
        c << "((" << b.lo[0] << "," << b.lo[1] << ") (" << b.hi[0] << "," << b.hi[1] << ") (0,0))\n";
    c << ")\n" << boxes.size() << "\n";
    for (size_t k = 0; k < boxes.size(); ++k) c << "FabOnDisk: Cell_D_00000 " << offsets[k] << "\n";
    c << "\n" << boxes.size() << "," << ncomp << "\n";
    for (size_t k = 0; k < boxes.size(); ++k) {
        for (int n = 0; n < ncomp; ++n) c << mins[k * ncomp + n] << ",";
        c << "\n";
    }
    c << "\n" << boxes.size() << "," << ncomp << "\n";
    for (size_t k = 0; k < boxes.size(); ++k) {
        for (int n = 0; n < ncomp; ++n) c << maxs[k * ncomp + n] << ",";
        c << "\n";
    }
    return (bool)c;
}

// WriteMultiLevelPlotfile (ZoomyAmr.cpp:880): Header for every level (level
// domains, steps, cell sizes, per-level box extents), then Level_<l>/.
inline bool write_plotfile_levels(const std::string& path, const std::vector<std::string>& names, double time,
                                  const double plo[2], const double phi[2], int ref_ratio, int step,
                                  const std::vector<PlotLevel>& levels) {
    const int ncomp = (int)names.size();
    const int finest = (int)levels.size() - 1;
    if (!make_dir(path)) return false;
    {
        std::ofstream h(path + "/Header");
        if (!h) return false;
        h << std::setprecision(17);
        h << "HyperCLaw-V1.1\n" << ncomp << "\n";
        for (auto& n : names) h << n << "\n";
        h << 2 << "\n" << time << "\n" << finest << "\n";
        h << plo[0] << " " << plo[1] << "\n" << phi[0] << " " << phi[1] << "\n";
        for (int l = 0; l < finest; ++l) h << ref_ratio << (l + 1 < finest ? " " : "");
        h << "\n";
        for (int l = 0; l <= finest; ++l)
            h << "((0,0) (" << levels[l].nx - 1 << "," << levels[l].ny - 1 << ") (0,0))" << (l < finest ? " " : "");
        h << "\n";
        for (int l = 0; l <= finest; ++l) h << step << (l < finest ? " " : "");
        h << "\n";
        for (int l = 0; l <= finest; ++l) h << levels[l].dx[0] << " " << levels[l].dx[1] << "\n";
        h << 0 << "\n" << 0 << "\n";
        for (int l = 0; l <= finest; ++l) {
            const PlotLevel& L = levels[l];
            h << l << " " << L.boxes.size() << " " << time << "\n" << step << "\n";
            for (auto& b : L.boxes) {
                h << plo[0] + b.lo[0] * L.dx[0] << " " << plo[0] + (b.hi[0] + 1) * L.dx[0] << "\n";
                h << plo[1] + b.lo[1] * L.dx[1] << " " << plo[1] + (b.hi[1] + 1) * L.dx[1] << "\n";
            }
            h << "Level_" << l << "/Cell\n";
        }
    }
    for (int l = 0; l <= finest; ++l)
        if (!write_plot_level(path, l, ncomp, levels[l].boxes)) return false;
    return true;
}

inline bool write_plotfile_level0(const std::string& path, const std::vector<std::string>& names, double time,
                                  const double plo[2], const double phi[2], int nx, int ny, const double dx[2],
                                  int step, const std::vector<PlotBox>& boxes) {
    std::vector<PlotLevel> lv(1);
    lv[0].nx = nx;
    lv[0].ny = ny;
    lv[0].dx[0] = dx[0];
    lv[0].dx[1] = dx[1];
    lv[0].boxes = boxes;
    return write_plotfile_levels(path, names, time, plo, phi, 2, step, lv);
}

}  // namespace zamr
