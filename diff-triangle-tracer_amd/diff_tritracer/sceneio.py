"""Scene files: triangle-splat scenes saved and loaded as binary little-endian PLY, with the
byte unpacking / packing done on the GPU (tt_scene_unpack, tt_scene_pack).

The reference keeps scenes as torch tensors handed to its autograd Function
(diff_tritracer/__init__.py:54-139: vertices [3P, 3], num_points_per_triangle,
cumsum_of_points_per_triangle, opacities, sigma, shs [P, M, 3]) and has no file format.
The format written here is a mesh-compatible PLY (viewable as a triangle mesh):

    element vertex 3P   float x, y, z
    element face P      list uchar int vertex_indices, float opacity, float sigma,
                        float f_dc_0..2, float f_rest_0..3(M-1)-1

with SH coefficient j >= 1 of channel c at f_rest_{c (M-1) + j - 1}. The loader also takes
shared vertices, extra vertex / face properties and any property order, as long as the
file is binary little endian, properties are scalars (besides vertex_indices) and the
indices are 32-bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi

_SIZES = {"char": 1, "uchar": 1, "int8": 1, "uint8": 1, "short": 2, "ushort": 2, "int16": 2, "uint16": 2,
          "int": 4, "uint": 4, "int32": 4, "uint32": 4, "float": 4, "float32": 4, "double": 8, "float64": 8}


def read_header(f):
    """Parse a PLY header from a binary file object positioned at its start. Returns
    (format, elements) with elements = [(name, count, [(prop, type) or (prop, (count_type,
    item_type))])]; the file is left at the first body byte."""
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements = None, []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        w = line.decode("ascii", "replace").split()
        if not w or w[0] in ("comment", "obj_info"):
            continue
        if w[0] == "end_header":
            return fmt, elements
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append((w[1], int(w[2]), []))
        elif w[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if w[1] == "list":
                elements[-1][2].append((w[4], (w[2], w[3])))
            else:
                elements[-1][2].append((w[2], w[1]))


def _layout(props, element):
    """Byte offsets of the scalar properties of one element and its row size."""
    offs, o = {}, 0
    for name, typ in props:
        if isinstance(typ, tuple):
            cnt, item = typ
            if name not in ("vertex_indices", "vertex_index") or _SIZES.get(cnt) != 1 or _SIZES.get(item) != 4:
                raise ValueError(f"{element}: only a 'list uchar int vertex_indices' list property is supported")
            offs[name] = o
            o += 1 + 3 * 4  # fixed-size rows: three indices
        else:
            if typ not in _SIZES:
                raise ValueError(f"{element}: unknown property type {typ}")
            offs[name] = (o, typ)
            o += _SIZES[typ]
    return offs, o


def load_ply(path, device="cuda", allow_bad=False):
    """Read a scene file into the tracer's layout on `device`: dict(vertices [3P, 3],
    npts [P], cumsum [P], opacity [P, 1], sigma [P, 1], shs [P, M, 3], M)."""
    with open(path, "rb") as f:
        fmt, elements = read_header(f)
        if fmt != "binary_little_endian":
            raise ValueError(f"{path}: format {fmt} is not supported (binary_little_endian only)")
        el = {name: (count, props) for name, count, props in elements}
        if "vertex" not in el or "face" not in el:
            raise ValueError(f"{path}: needs a vertex and a face element")
        bodies = {}
        for name, count, props in elements:  # elements are stored in header order
            _, size = _layout(props, name)
            if any(isinstance(t, tuple) for _, t in props) and name != "face":
                raise ValueError(f"{path}: list properties are supported on the face element only")
            data = f.read(count * size)
            if len(data) != count * size:
                raise ValueError(f"{path}: truncated {name} element")
            bodies[name] = (data, size, _layout(props, name)[0])
    V, P = el["vertex"][0], el["face"][0]
    vbytes, vstride, voffs = bodies["vertex"]
    fbytes, fstride, foffs = bodies["face"]
    for c in "xyz":
        if c not in voffs or voffs[c][1] not in ("float", "float32"):
            raise ValueError(f"{path}: vertex property {c} (float) missing")
    idx_name = "vertex_indices" if "vertex_indices" in foffs else "vertex_index"
    if idx_name not in foffs:
        raise ValueError(f"{path}: face element has no vertex_indices list")

    def fo(name):
        if name not in foffs:
            return -1
        o, typ = foffs[name]
        if typ not in ("float", "float32"):
            raise ValueError(f"{path}: face property {name} must be float")
        return o

    n_rest = sum(1 for k in foffs if k.startswith("f_rest_"))
    if n_rest % 3:
        raise ValueError(f"{path}: {n_rest} f_rest properties (not 3 per coefficient)")
    M = 1 + n_rest // 3
    if M > capi.SCENE_MAX_COEFFS:
        raise ValueError(f"{path}: {M} SH coefficients per channel (at most {capi.SCENE_MAX_COEFFS})")
    a = capi.SceneUnpackArgs()
    a.P, a.V, a.face_stride, a.M = P, V, fstride, M
    a.count_offset, a.index_offset = foffs[idx_name], foffs[idx_name] + 1
    a.opacity_offset, a.sigma_offset = fo("opacity"), fo("sigma")
    for k in range(capi.SCENE_MAX_COEFFS * 3):
        a.sh_offset[k] = -1
    for c in range(3):
        a.sh_offset[c] = fo(f"f_dc_{c}")
        for j in range(1, M):
            a.sh_offset[3 * j + c] = fo(f"f_rest_{c * (M - 1) + j - 1}")
    a.vert_stride = vstride
    for c, name in enumerate("xyz"):
        a.xyz_offset[c] = voffs[name][0]
    dev = torch.device(device)
    face_dev = torch.from_numpy(np.frombuffer(fbytes, np.uint8).copy()).to(dev) if P else None
    vert_dev = torch.from_numpy(np.frombuffer(vbytes, np.uint8).copy()).to(dev) if V else None
    out = dict(vertices=torch.empty(3 * P, 3, device=dev), npts=torch.empty(P, dtype=torch.int32, device=dev),
               cumsum=torch.empty(P, dtype=torch.int32, device=dev), opacity=torch.empty(P, 1, device=dev),
               sigma=torch.empty(P, 1, device=dev), shs=torch.empty(P, M, 3, device=dev), M=M)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    a.face_dev, a.vert_dev = capi.ptr(face_dev), capi.ptr(vert_dev)
    a.vertices_out_dev, a.npts_out_dev, a.cumsum_out_dev = (capi.ptr(out[k]) for k in ("vertices", "npts", "cumsum"))
    a.opacity_out_dev, a.sigma_out_dev, a.shs_out_dev = (capi.ptr(out[k]) for k in ("opacity", "sigma", "shs"))
    a.bad_dev = capi.ptr(bad)
    capi.check(capi.lib().tt_scene_unpack(C.byref(a), torch.cuda.current_stream(dev).cuda_stream), "tt_scene_unpack")
    nbad = int(bad.item())
    if nbad and not allow_bad:
        raise ValueError(f"{path}: {nbad} faces have a list count other than 3 or an index outside [0, {V})")
    out["bad_faces"] = nbad
    return out


def header_text(P: int, M: int) -> bytes:
    lines = ["ply", "format binary_little_endian 1.0",
             "comment triangle splats: per-face opacity, sigma, SH (f_dc_c, f_rest_{c (M-1) + j - 1})",
             f"element vertex {3 * P}", "property float x", "property float y", "property float z",
             f"element face {P}", "property list uchar int vertex_indices", "property float opacity",
             "property float sigma"]
    lines += [f"property float f_dc_{c}" for c in range(3)]
    lines += [f"property float f_rest_{k}" for k in range(3 * (M - 1))]
    lines.append("end_header")
    return ("\n".join(lines) + "\n").encode("ascii")


def pack(vertices, cumsum, opacity, sigma, shs):
    """Device tensors -> (vertex rows float32 [3P, 3], face rows uint8 [P * row bytes]) on the device."""
    P = int(cumsum.numel())
    M = int(shs.shape[1]) if shs.dim() == 3 else int(shs.numel() // max(P, 1) // 3)
    for t, n in ((vertices, "vertices"), (opacity, "opacity"), (sigma, "sigma"), (shs, "shs")):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise ValueError(f"{n} must be a contiguous float32 GPU tensor")
    if not (cumsum.is_cuda and cumsum.dtype == torch.int32 and cumsum.is_contiguous()):
        raise ValueError("cumsum must be a contiguous int32 GPU tensor")
    dev = vertices.device
    vert_out = torch.empty(3 * P, 3, device=dev)
    face_out = torch.empty(P * capi.scene_face_bytes(M), dtype=torch.uint8, device=dev)
    a = capi.ScenePackArgs(P, M, capi.ptr(vertices), capi.ptr(cumsum), capi.ptr(opacity), capi.ptr(sigma),
                           capi.ptr(shs), capi.ptr(face_out), capi.ptr(vert_out))
    capi.check(capi.lib().tt_scene_pack(C.byref(a), torch.cuda.current_stream(dev).cuda_stream), "tt_scene_pack")
    return vert_out, face_out, M


def save_ply(path, vertices, cumsum, opacity, sigma, shs):
    """Write a scene (tracer-layout device tensors) to `path`."""
    P = int(cumsum.numel())
    vert_out, face_out, M = pack(vertices, cumsum, opacity, sigma, shs)
    with open(path, "wb") as f:
        f.write(header_text(P, M))
        f.write(vert_out.cpu().numpy().tobytes())
        f.write(face_out.cpu().numpy().tobytes())
