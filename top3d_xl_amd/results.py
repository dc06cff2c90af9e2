"""Result and trajectory files of a design run, in the reference's formats.

* iters_Target.dat  : PCG iterations per design iteration, '%d\\n' (193-195, 410-412)
* c_Target.dat      : compliance history, '%30.16e\\n' (196-198, 413-415)
* sharp_Target.dat  : sharpness history, '%30.16e\\n' (199-201, 416-418)
* timing_Target.dat : total time per design iteration, '%16.6e\\n' (202-204, 419-421)
* DesignVolume.stl  : closed 0.5 iso-surface of that volume, flipped along y
                      (206-222, 572-577; see isosurface.py)
* DesignVolume.nii  : xPhys scattered into the padded (resY, resX, resZ) grid,
                      zeros elsewhere, as a NIfTI-1 float64 volume
                      (IO_ExportDesignInVolume_nii, 564-570; MATLAB's
                      niftiwrite default header values are not reproduced
                      beyond dimensions, data type and voxel size 1)
"""
from __future__ import annotations

import os
import struct

import numpy as np


def write_target_files(out_dir, result):
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "iters_Target.dat"), "w") as f:
        f.writelines(f"{int(v)}\n" for v in result["lssIts"])
    with open(os.path.join(out_dir, "c_Target.dat"), "w") as f:
        f.writelines(f"{v:30.16e}\n" for v in result["cHist"])
    with open(os.path.join(out_dir, "sharp_Target.dat"), "w") as f:
        f.writelines(f"{v:30.16e}\n" for v in result["sharpHist"])
    with open(os.path.join(out_dir, "timing_Target.dat"), "w") as f:
        f.writelines(f"{v:16.6e}\n" for v in np.asarray(result["tHist"])[:, -1])


def cell_map_from_voxels(voxels, res):
    """eleMapBack (1684-1689): padded column-major index of each solid voxel,
    in the compact cell order (ascending padded index)."""
    resY, resX, resZ = res
    iy, ix, iz = np.nonzero(np.asarray(voxels, dtype=bool))
    idx = iy + resY * (ix + resX * iz.astype(np.int64))
    return np.sort(idx)


def padded_volume(xPhys, cell_map, res):
    """xPhys (compact cells) -> (resY, resX, resZ) array; cell_map holds the
    padded-grid column-major index of every compact cell (eleMapBack)."""
    resY, resX, resZ = res
    V = np.zeros(resY * resX * resZ)
    V[np.asarray(cell_map)] = xPhys
    return V.reshape((resY, resX, resZ), order="F")


def write_nii(path, vol):
    """Single-file NIfTI-1 (.nii), float64, unit voxel size, data column-major."""
    vol = np.asarray(vol, dtype="<f8")
    dims = list(vol.shape) + [1] * (7 - vol.ndim)
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)                                 # sizeof_hdr
    struct.pack_into("<8h", hdr, 40, vol.ndim, *dims)                   # dim[0..7]
    struct.pack_into("<hh", hdr, 70, 64, 64)                            # datatype DOUBLE, bitpix
    struct.pack_into("<8f", hdr, 76, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)  # pixdim
    struct.pack_into("<f", hdr, 108, 352.0)                             # vox_offset
    struct.pack_into("<f", hdr, 112, 1.0)                               # scl_slope
    hdr[344:348] = b"n+1\0"                                             # magic
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(vol.reshape(-1, order="F").tobytes())


def read_nii(path):
    with open(path, "rb") as f:
        raw = f.read()
    ndim = struct.unpack_from("<h", raw, 40)[0]
    dims = struct.unpack_from("<8h", raw, 40)[1:1 + ndim]
    dtype, bitpix = struct.unpack_from("<hh", raw, 70)
    if dtype != 64 or bitpix != 64 or raw[344:347] != b"n+1":
        raise ValueError("not a float64 NIfTI-1 file")
    off = int(struct.unpack_from("<f", raw, 108)[0])
    n = int(np.prod(dims))
    return np.frombuffer(raw, dtype="<f8", count=n, offset=off).reshape(dims, order="F")


def write_design_outputs(out_dir, result, voxels, res):
    """Everything a design run leaves in outPath: the *_Target.dat files
    (193-204), DesignVolume.nii (185-186, 564-570) and DesignVolume.stl
    (206-221).  res is the padded (resY, resX, resZ) cell grid."""
    from .isosurface import write_design_stl
    write_target_files(out_dir, result)
    cm = cell_map_from_voxels(voxels, res)
    write_nii(os.path.join(out_dir, "DesignVolume.nii"), padded_volume(result["xPhys"], cm, res))
    write_design_stl(os.path.join(out_dir, "DesignVolume.stl"), result["xPhys"], cm, res)
