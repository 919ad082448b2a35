// VTK ImageData output of cell fields, the writer of the reference's test
// programs (c++/test/unbounded3d/sources/main.cpp:329-455): one .vti piece
// per rank with 3-component Float64 CellData arrays, and a .pvti file that
// lists the pieces.  One rank here, so one piece covering the grid.
// Host code only.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ps_capi.h"

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const std::string& path) : f(std::fopen(path.c_str(), "wb")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

std::string extent(const int n[3])
{
    char b[128];
    std::snprintf(b, sizeof b, "  0  %d  0  %d  0  %d", n[0], n[1], n[2]);
    return b;
}

std::string triple(const double v[3])
{
    char b[128];
    std::snprintf(b, sizeof b, "  %.8g  %.8g  %.8g", v[0], v[1], v[2]);
    return b;
}

}  // namespace

extern "C" int ps_write_vti(const char* dir, const char* stem, const int ncell[3], const double origin[3],
                            const double dx[3], int nfield, const char* const* names, const double* const* comps,
                            int binary)
{
    if (!dir || !stem || !ncell || !origin || !dx || nfield < 0 || (nfield > 0 && (!names || !comps)))
        return PS_EINVAL;
    for (int d = 0; d < 3; ++d)
        if (ncell[d] < 1) return PS_EINVAL;
    for (int a = 0; a < nfield; ++a)
        if (!names[a]) return PS_EINVAL;
    const long n3 = (long)ncell[0] * ncell[1] * ncell[2];
    const std::string base = std::string(dir) + "/" + stem;
    const std::string piece = std::string(stem) + "_P00.vti";
    {
        File v(std::string(dir) + "/" + piece);
        if (!v.f) return PS_EINVAL;
        std::fprintf(v.f, "<?xml version='1.0'?>\n");
        std::fprintf(v.f, "<VTKFile type='ImageData' version='0.1' byte_order='LittleEndian'%s>\n",
                     binary ? " header_type='UInt64'" : "");
        std::fprintf(v.f, "  <ImageData WholeExtent='%s' Ghostlevel='0' Origin='%s' Spacing='%s'>\n",
                     extent(ncell).c_str(), triple(origin).c_str(), triple(dx).c_str());
        std::fprintf(v.f, "    <Piece Extent='%s'>\n", extent(ncell).c_str());
        std::fprintf(v.f, "      <PointData>\n      </PointData>\n      <CellData>\n");
        const uint64_t bytes = (uint64_t)n3 * 3 * sizeof(double);
        for (int a = 0; a < nfield; ++a) {
            const double* c[3] = {comps[3 * a], comps[3 * a + 1], comps[3 * a + 2]};
            if (binary) {
                std::fprintf(v.f,
                             "        <DataArray type='Float64' Name='%s' NumberOfComponents='3' format='appended' "
                             "offset='%llu'/>\n",
                             names[a], (unsigned long long)(a * (bytes + 8)));
                continue;
            }
            std::fprintf(v.f, "        <DataArray type='Float64' Name='%s' NumberOfComponents='3'  format='ascii'>\n",
                         names[a]);
            for (long i = 0; i < n3; ++i)
                std::fprintf(v.f, "%17.8e%17.8e%17.8e", c[0] ? c[0][i] : 0.0, c[1] ? c[1][i] : 0.0,
                             c[2] ? c[2][i] : 0.0);
            std::fprintf(v.f, "\n        </DataArray>\n");
        }
        std::fprintf(v.f, "      </CellData>\n    </Piece>\n  </ImageData>\n");
        if (binary) {
            // raw appended block: per array a UInt64 byte count, then the
            // interleaved components
            std::fprintf(v.f, "  <AppendedData encoding='raw'>\n   _");
            std::vector<double> buf;
            for (int a = 0; a < nfield; ++a) {
                const double* c[3] = {comps[3 * a], comps[3 * a + 1], comps[3 * a + 2]};
                std::fwrite(&bytes, sizeof bytes, 1, v.f);
                const long chunk = 1 << 16;
                buf.resize(3 * chunk);
                for (long i0 = 0; i0 < n3; i0 += chunk) {
                    const long m = (n3 - i0 < chunk) ? n3 - i0 : chunk;
                    for (long i = 0; i < m; ++i)
                        for (int d = 0; d < 3; ++d) buf[3 * i + d] = c[d] ? c[d][i0 + i] : 0.0;
                    if (std::fwrite(buf.data(), sizeof(double), 3 * m, v.f) != (size_t)(3 * m)) return PS_EINVAL;
                }
            }
            std::fprintf(v.f, "\n  </AppendedData>\n");
        }
        std::fprintf(v.f, "</VTKFile>\n");
        if (std::ferror(v.f)) return PS_EINVAL;
    }
    File p(base + ".pvti");
    if (!p.f) return PS_EINVAL;
    std::fprintf(p.f, "<?xml version='1.0'?>\n");
    std::fprintf(p.f, "<VTKFile type='PImageData' version='0.1' byte_order='LittleEndian'>\n");
    std::fprintf(p.f, "<PImageData WholeExtent='%s' Ghostlevel='0' Origin='%s' Spacing='%s'>\n",
                 extent(ncell).c_str(), triple(origin).c_str(), triple(dx).c_str());
    std::fprintf(p.f, "  <PCellData Vectors='output'>\n");
    for (int a = 0; a < nfield; ++a)
        std::fprintf(p.f, "    <PDataArray type='Float64' Name='%s' NumberOfComponents='3' format='appended' "
                          "offset='0'/>\n",
                     names[a]);
    std::fprintf(p.f, "  </PCellData>\n");
    std::fprintf(p.f, "  <Piece Extent='%s' Source='./%s'/>\n", extent(ncell).c_str(), piece.c_str());
    std::fprintf(p.f, "</PImageData>\n</VTKFile>\n");
    return std::ferror(p.f) ? PS_EINVAL : PS_OK;
}
