// Writes ConBlock_r0_<snap>.vtp / ConBlock_<snap>.pvtp (include/alens/ConBlockVTK.hpp) for the
// alens_con_block records stored raw in argv[1]: usage conblock_write blocks.bin folder snap
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "alens/ConBlockVTK.hpp"

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    std::vector<alens_con_block> b;
    alens_con_block one;
    while (std::fread(&one, sizeof one, 1, f) == 1) b.push_back(one);
    std::fclose(f);
    if (!alens::vtk::writeConBlocks(argv[2], std::atoi(argv[3]), b)) return 4;
    std::printf("%zu\n", b.size());
    return 0;
}
