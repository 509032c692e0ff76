"""Output files of TOP_GPU.m in the reference's formats (TOP_GPU.m:294-326, 539-571).

    c_Target.dat          compliance history        '%30.16e\\n'
    iters_Target.dat      PCG iterations per solve  '%d\\n'
    sharp_Target.dat      sharpness history         '%30.16e\\n'
    timing_Target.dat     per-iteration total time  '%16.6e\\n'
    solvetiming_Target.dat  per-iteration CG time   '%16.6e\\n'
    DesignVolume.nii      xPhys on the padded voxel grid (NIfTI-1, float64), as niftiwrite(V)
"""
import os
import struct

import numpy as np


def write_histories(path, out):
    os.makedirs(path, exist_ok=True)

    def w(name, vals, fmt):
        with open(os.path.join(path, name), 'w') as f:
            for v in vals:
                f.write(fmt % v)

    w('iters_Target.dat', out['lssIts'], '%d\n')
    w('c_Target.dat', out['cHist'], '%30.16e\n')
    w('sharp_Target.dat', out['sharpHist'], '%30.16e\n')
    w('timing_Target.dat', [t[-1] for t in out['tHist']], '%16.6e\n')
    w('solvetiming_Target.dat', [t[1] for t in out['tHist']], '%16.6e\n')
    model = out.get('model')
    if model is not None:
        write_design_nii(os.path.join(path, 'DesignVolume.nii'), model, out['xPhys'])


def read_history(path, name):
    with open(os.path.join(path, name)) as f:
        return [float(t) for t in f.read().split()]


def design_volume(model, xPhys):
    """IO_ExportDesignInVolume_nii's V: densities on the padded voxel grid (super-element: the
    num-times refined grid), zero for void voxels."""
    info = model.level_info(1)
    ry, rx, rz = info['resY'], info['resX'], info['resZ']
    x = np.asarray(xPhys.cpu().numpy() if hasattr(xPhys, 'cpu') else xPhys, dtype=np.float64).reshape(-1)
    back = model.get_map(1, 'eleMapBack').astype(np.int64) - 1
    if not model.super_element:
        V = np.zeros(rx * ry * rz)
        V[back] = x
        return V.reshape((ry, rx, rz), order='F')
    up = 4
    V = np.zeros((ry * up, rx * up, rz * up))
    blk = x.reshape(up ** 3, -1, order='F')
    for t, e in enumerate(back):
        iy, ix, iz = e % ry, (e // ry) % rx, e // (ry * rx)
        V[iy * up:(iy + 1) * up, ix * up:(ix + 1) * up, iz * up:(iz + 1) * up] = \
            blk[:, t].reshape((up, up, up), order='F')
    return V


def write_nifti1(fileName, V):
    """Minimal NIfTI-1 single file (.nii): 348-byte header, 4-byte extension flag, float64 data."""
    V = np.asarray(V, dtype=np.float64)
    dims = list(V.shape) + [1] * (7 - V.ndim)
    hdr = bytearray(348)
    struct.pack_into('<i', hdr, 0, 348)                     # sizeof_hdr
    struct.pack_into('<8h', hdr, 40, V.ndim, *dims)         # dim
    struct.pack_into('<h', hdr, 70, 64)                     # datatype: DT_FLOAT64
    struct.pack_into('<h', hdr, 72, 64)                     # bitpix
    struct.pack_into('<8f', hdr, 76, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)  # pixdim
    struct.pack_into('<f', hdr, 108, 352.0)                 # vox_offset
    struct.pack_into('<f', hdr, 112, 1.0)                   # scl_slope
    hdr[344:348] = b'n+1\0'                                 # magic
    with open(fileName, 'wb') as f:
        f.write(bytes(hdr))
        f.write(b'\0\0\0\0')
        f.write(V.reshape(-1, order='F').tobytes())


def read_nifti1(fileName):
    raw = open(fileName, 'rb').read()
    ndim, *dims = struct.unpack_from('<8h', raw, 40)
    off = int(struct.unpack_from('<f', raw, 108)[0])
    shape = tuple(dims[:ndim])
    return np.frombuffer(raw[off:], dtype=np.float64).reshape(shape, order='F')


def write_design_nii(fileName, model, xPhys):
    write_nifti1(fileName, design_volume(model, xPhys))
