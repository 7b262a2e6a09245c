// Test helper: reads a legacy STRUCTURED_POINTS fluid mask with the C++
// front-end (lbm::io::read_vtk_mask, include/lbm_hip/io.hpp) and prints its
// dimensions, origin, spacing, fluid count and the mask as one line of 0/1.
#include <cstdio>

#include "lbm_hip/io.hpp"

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    try {
        const lbm::io::VtkMask m = lbm::io::read_vtk_mask(argv[1]);
        size_t nf = 0;
        for (uint8_t v : m.fluid) nf += v;
        std::printf("dims %zu %zu %zu\n", m.xl, m.yl, m.zl);
        std::printf("origin %.9g %.9g %.9g\n", m.origin[0], m.origin[1], m.origin[2]);
        std::printf("spacing %.9g %.9g %.9g\n", m.spacing[0], m.spacing[1], m.spacing[2]);
        std::printf("fluid %zu\n", nf);
        std::string s(m.fluid.size(), '0');
        for (size_t i = 0; i < m.fluid.size(); ++i) s[i] = m.fluid[i] ? '1' : '0';
        std::printf("mask %s\n", s.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
