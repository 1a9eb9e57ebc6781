// ConBlockVTK.hpp -- constraint-block snapshots as VTK XML PolyData (ConBlock_*.pvtp / .vtp).
//
// The reference writes, per snapshot, the collision and protein constraint blocks of each rank to
// ConBlock_r<rank>_<snap>.vtp and an index ConBlock_<snap>.pvtp beside it (README.md:136-148,
// docs/source/visualization.md:7-11); Run/result/Result2PVD.py:26 collects the .pvtp files into
// ConBlockpvtp.pvd.  The arrays are those the reference's ParaView state reads from these files
// (Examples/MixMotorSliding/MixMotorSliding.pvsm, reader of ConBlockpvtp.pvd):
//   points  : two per block, lab_i and lab_j (lab frame), joined by one line cell;
//   point data: gid, globalIndex (Int32), posIJ, normIJ (Float64 x 3) -- the I side on the first
//               point of the block, the J side on the second;
//   cell data : oneSide, bilateral (Int32), delta0, gamma, kappa (Float64), Stress (Float64 x 9).
// Data arrays are written inline as base64 "binary" with UInt64 byte-count headers (the header and
// the data encoded separately, as VTK's own writer does), little-endian.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../alens_capi.h"

namespace alens {
namespace vtk {

inline std::string base64(const unsigned char *p, size_t n) {
    static const char *tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve(4 * ((n + 2) / 3));
    size_t i = 0;
    for (; i + 3 <= n; i += 3) {
        const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | (uint32_t)p[i + 2];
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += tab[(v >> 6) & 63];
        out += tab[v & 63];
    }
    if (i < n) {
        const uint32_t v = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0u);
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += i + 1 < n ? tab[(v >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

template <typename T>
inline void data_array(std::FILE *f, const char *type, const char *name, int ncomp, const std::vector<T> &v) {
    const uint64_t nbytes = (uint64_t)(v.size() * sizeof(T));
    unsigned char hdr[8];
    std::memcpy(hdr, &nbytes, 8);  // little-endian hosts only (x86-64)
    std::fprintf(f, "        <DataArray type=\"%s\"", type);
    if (name) std::fprintf(f, " Name=\"%s\"", name);
    std::fprintf(f, " NumberOfComponents=\"%d\" format=\"binary\">\n          %s%s\n        </DataArray>\n", ncomp,
                 base64(hdr, 8).c_str(), base64((const unsigned char *)v.data(), (size_t)nbytes).c_str());
}

// Writes folder/ConBlock_r0_<snap>.vtp and folder/ConBlock_<snap>.pvtp (one piece: one GPU holds
// every block).  Returns false if a file cannot be written.
inline bool writeConBlocks(const std::string &folder, int snap, const std::vector<alens_con_block> &blocks) {
    const size_t m = blocks.size();
    std::vector<double> pts(6 * m), pos(6 * m), nrm(6 * m), d0(m), gam(m), kap(m), stress(9 * m);
    std::vector<int32_t> gid(2 * m), gidx(2 * m), one(m), bil(m), conn(2 * m), offs(m);
    for (size_t k = 0; k < m; ++k) {
        const alens_con_block &b = blocks[k];
        for (int a = 0; a < 3; ++a) {
            pts[6 * k + a] = b.lab_i[a];
            pts[6 * k + 3 + a] = b.lab_j[a];
            pos[6 * k + a] = b.pos_i[a];
            pos[6 * k + 3 + a] = b.pos_j[a];
            nrm[6 * k + a] = b.norm_i[a];
            nrm[6 * k + 3 + a] = b.norm_j[a];
        }
        gid[2 * k] = b.gid_i;
        gid[2 * k + 1] = b.gid_j;
        gidx[2 * k] = b.gidx_i;
        gidx[2 * k + 1] = b.gidx_j;
        one[k] = b.one_side;
        bil[k] = b.bilateral;
        d0[k] = b.delta0;
        gam[k] = b.gamma;
        kap[k] = b.kappa;
        std::memcpy(&stress[9 * k], b.stress, sizeof b.stress);
        conn[2 * k] = (int32_t)(2 * k);
        conn[2 * k + 1] = (int32_t)(2 * k + 1);
        offs[k] = (int32_t)(2 * k + 2);
    }
    const std::string piece = "ConBlock_r0_" + std::to_string(snap) + ".vtp";
    std::FILE *f = std::fopen((folder + "/" + piece).c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" "
                    "header_type=\"UInt64\">\n  <PolyData>\n    <Piece NumberOfPoints=\"%zu\" NumberOfVerts=\"0\" "
                    "NumberOfLines=\"%zu\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n      <Points>\n",
                 2 * m, m);
    data_array(f, "Float64", nullptr, 3, pts);
    std::fprintf(f, "      </Points>\n      <PointData>\n");
    data_array(f, "Int32", "gid", 1, gid);
    data_array(f, "Int32", "globalIndex", 1, gidx);
    data_array(f, "Float64", "posIJ", 3, pos);
    data_array(f, "Float64", "normIJ", 3, nrm);
    std::fprintf(f, "      </PointData>\n      <CellData>\n");
    data_array(f, "Int32", "oneSide", 1, one);
    data_array(f, "Int32", "bilateral", 1, bil);
    data_array(f, "Float64", "delta0", 1, d0);
    data_array(f, "Float64", "gamma", 1, gam);
    data_array(f, "Float64", "kappa", 1, kap);
    data_array(f, "Float64", "Stress", 9, stress);
    std::fprintf(f, "      </CellData>\n      <Lines>\n");
    data_array(f, "Int32", "connectivity", 1, conn);
    data_array(f, "Int32", "offsets", 1, offs);
    std::fprintf(f, "      </Lines>\n    </Piece>\n  </PolyData>\n</VTKFile>\n");
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) return false;

    f = std::fopen((folder + "/ConBlock_" + std::to_string(snap) + ".pvtp").c_str(), "w");
    if (!f) return false;
    std::fprintf(f,
                 "<?xml version=\"1.0\"?>\n<VTKFile type=\"PPolyData\" version=\"1.0\" byte_order=\"LittleEndian\" "
                 "header_type=\"UInt64\">\n  <PPolyData GhostLevel=\"0\">\n"
                 "    <PPoints>\n      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n    </PPoints>\n"
                 "    <PPointData>\n"
                 "      <PDataArray type=\"Int32\" Name=\"gid\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Int32\" Name=\"globalIndex\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"posIJ\" NumberOfComponents=\"3\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"normIJ\" NumberOfComponents=\"3\"/>\n"
                 "    </PPointData>\n    <PCellData>\n"
                 "      <PDataArray type=\"Int32\" Name=\"oneSide\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Int32\" Name=\"bilateral\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"delta0\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"gamma\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"kappa\" NumberOfComponents=\"1\"/>\n"
                 "      <PDataArray type=\"Float64\" Name=\"Stress\" NumberOfComponents=\"9\"/>\n"
                 "    </PCellData>\n    <Piece Source=\"%s\"/>\n  </PPolyData>\n</VTKFile>\n",
                 piece.c_str());
    const bool ok2 = std::ferror(f) == 0;
    return std::fclose(f) == 0 && ok2;
}

}  // namespace vtk
}  // namespace alens
