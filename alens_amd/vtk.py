"""Constraint-block snapshots as VTK XML PolyData (ConBlock_r0_<snap>.vtp + ConBlock_<snap>.pvtp).

The Python mirror of include/alens/ConBlockVTK.hpp (same layout, same bytes): the reference writes
these files per snapshot (README.md:136-148, docs/source/visualization.md:7-11) and
Run/result/Result2PVD.py:26 indexes them.  Arrays, as the reference's ParaView state reads them
(Examples/MixMotorSliding/MixMotorSliding.pvsm): two points per block (lab_i, lab_j) joined by a line
cell; point data gid, globalIndex (Int32), posIJ, normIJ (Float64 x 3); cell data oneSide, bilateral
(Int32), delta0, gamma, kappa (Float64), Stress (Float64 x 9).  Inline base64 "binary" arrays with
UInt64 byte-count headers, header and data encoded separately.
"""
import base64
import os

import numpy as np


def _array(type_, name, ncomp, data):
    raw = np.ascontiguousarray(data).tobytes()
    hdr = np.uint64(len(raw)).tobytes()
    nm = f' Name="{name}"' if name else ""
    return (f'        <DataArray type="{type_}"{nm} NumberOfComponents="{ncomp}" format="binary">\n'
            f'          {base64.b64encode(hdr).decode()}{base64.b64encode(raw).decode()}\n'
            f'        </DataArray>\n')


def blocks_to_arrays(blocks):
    """Structured array (or list) of alens_con_block records -> the per-point / per-cell arrays."""
    b = np.asarray(blocks)
    m = b.shape[0]
    f64 = lambda key: np.asarray(b[key], np.float64).reshape(m, int(np.prod(b.dtype[key].shape, dtype=int)))
    pts = np.stack([f64("lab_i"), f64("lab_j")], axis=1).reshape(-1, 3)
    pos = np.stack([f64("pos_i"), f64("pos_j")], axis=1).reshape(-1, 3)
    nrm = np.stack([f64("norm_i"), f64("norm_j")], axis=1).reshape(-1, 3)
    gid = np.stack([b["gid_i"], b["gid_j"]], axis=1).reshape(-1).astype(np.int32)
    gidx = np.stack([b["gidx_i"], b["gidx_j"]], axis=1).reshape(-1).astype(np.int32)
    return dict(points=pts, gid=gid, globalIndex=gidx, posIJ=pos, normIJ=nrm,
                oneSide=np.asarray(b["one_side"], np.int32), bilateral=np.asarray(b["bilateral"], np.int32),
                delta0=f64("delta0").reshape(-1), gamma=f64("gamma").reshape(-1), kappa=f64("kappa").reshape(-1),
                Stress=f64("stress"))


def write_conblocks(folder, snap, blocks):
    """Write folder/ConBlock_r0_<snap>.vtp and folder/ConBlock_<snap>.pvtp; returns the .pvtp path."""
    a = blocks_to_arrays(blocks)
    m = a["delta0"].size
    conn = np.arange(2 * m, dtype=np.int32)
    offs = 2 * np.arange(1, m + 1, dtype=np.int32)
    piece = f"ConBlock_r0_{int(snap)}.vtp"
    parts = ['<?xml version="1.0"?>\n<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
             'header_type="UInt64">\n  <PolyData>\n'
             f'    <Piece NumberOfPoints="{2 * m}" NumberOfVerts="0" NumberOfLines="{m}" NumberOfStrips="0" '
             'NumberOfPolys="0">\n      <Points>\n',
             _array("Float64", None, 3, a["points"]), "      </Points>\n      <PointData>\n",
             _array("Int32", "gid", 1, a["gid"]), _array("Int32", "globalIndex", 1, a["globalIndex"]),
             _array("Float64", "posIJ", 3, a["posIJ"]), _array("Float64", "normIJ", 3, a["normIJ"]),
             "      </PointData>\n      <CellData>\n",
             _array("Int32", "oneSide", 1, a["oneSide"]), _array("Int32", "bilateral", 1, a["bilateral"]),
             _array("Float64", "delta0", 1, a["delta0"]), _array("Float64", "gamma", 1, a["gamma"]),
             _array("Float64", "kappa", 1, a["kappa"]), _array("Float64", "Stress", 9, a["Stress"]),
             "      </CellData>\n      <Lines>\n",
             _array("Int32", "connectivity", 1, conn), _array("Int32", "offsets", 1, offs),
             "      </Lines>\n    </Piece>\n  </PolyData>\n</VTKFile>\n"]
    with open(os.path.join(folder, piece), "w") as fh:
        fh.write("".join(parts))
    arr = lambda t, n, c: f'      <PDataArray type="{t}" Name="{n}" NumberOfComponents="{c}"/>\n'
    idx = ('<?xml version="1.0"?>\n<VTKFile type="PPolyData" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64">\n  <PPolyData GhostLevel="0">\n'
           '    <PPoints>\n      <PDataArray type="Float64" NumberOfComponents="3"/>\n    </PPoints>\n'
           '    <PPointData>\n' + arr("Int32", "gid", 1) + arr("Int32", "globalIndex", 1) + arr("Float64", "posIJ", 3)
           + arr("Float64", "normIJ", 3) + '    </PPointData>\n    <PCellData>\n' + arr("Int32", "oneSide", 1)
           + arr("Int32", "bilateral", 1) + arr("Float64", "delta0", 1) + arr("Float64", "gamma", 1)
           + arr("Float64", "kappa", 1) + arr("Float64", "Stress", 9)
           + f'    </PCellData>\n    <Piece Source="{piece}"/>\n  </PPolyData>\n</VTKFile>\n')
    path = os.path.join(folder, f"ConBlock_{int(snap)}.pvtp")
    with open(path, "w") as fh:
        fh.write(idx)
    return path
