"""Independent numpy decode of a glTF 2.0 file, used to check the C++ loader.

Follows the glTF 2.0 specification (accessors, buffer views with byteStride,
normalized integer components, node matrix / TRS, data: URIs and .bin files)
and the scene semantics of src/scene/mod.rs:106-295 (per-mesh concatenated
vertices with defaults, u32 indices, the first node subtree that holds the
mesh gives its global transform, the first perspective camera).  Written from
the specification in float64, not from the C++ code it checks.
"""
import base64
import json
import os
import struct

import numpy as np

_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def _load(path):
    if path.endswith(".glb"):
        raw = open(path, "rb").read()
        magic, _, length = struct.unpack_from("<III", raw, 0)
        assert magic == 0x46546C67
        off, doc, binchunk = 12, None, None
        while off < length:
            clen, ctype = struct.unpack_from("<II", raw, off)
            data = raw[off + 8:off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(data.decode())
            elif ctype == 0x004E4942:
                binchunk = data
            off += 8 + clen
        return doc, binchunk
    return json.load(open(path)), None


def _buffers(doc, path, binchunk):
    out = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            out.append(binchunk)
        elif uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            out.append(open(os.path.join(os.path.dirname(path), uri), "rb").read())
    return out


def _accessor(doc, bufs, i):
    a = doc["accessors"][i]
    dt = np.dtype(_COMP[a["componentType"]])
    n, nc = a["count"], _NCOMP[a["type"]]
    v = doc["bufferViews"][a["bufferView"]]
    data = bufs[v["buffer"]]
    base = v.get("byteOffset", 0) + a.get("byteOffset", 0)
    stride = v.get("byteStride", 0) or dt.itemsize * nc
    rows = [np.frombuffer(data, dt, nc, base + k * stride) for k in range(n)]
    arr = np.array(rows).reshape(n, nc)
    if a.get("normalized", False) and dt.kind in "iu":
        mx = float(np.iinfo(dt).max)
        arr = np.maximum(arr.astype(np.float64) / mx, -1.0)
    return arr


def _node_matrix(n):
    if "matrix" in n:
        return np.array(n["matrix"], np.float64).reshape(4, 4).T   # column-major
    t = np.array(n.get("translation", [0, 0, 0]), np.float64)
    x, y, z, w = n.get("rotation", [0, 0, 0, 1])
    s = np.array(n.get("scale", [1, 1, 1]), np.float64)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s[None, :]
    m[:3, 3] = t
    return m


def _global_transform(doc, mesh):
    nodes = doc.get("nodes", [])

    def walk(i, chain):
        chain = chain + [_node_matrix(nodes[i])]
        if nodes[i].get("mesh") == mesh:
            return chain
        for c in nodes[i].get("children", []):
            r = walk(c, chain)
            if r is not None:
                return r
        return None

    for i in range(len(nodes)):
        chain = walk(i, [])
        if chain is not None:
            g = np.eye(4)
            for m in chain:
                g = g @ m
            return g
    return np.eye(4)


def decode(path):
    """{"meshes": [{"vertices": (n,16) f64 [pos4 color4 normal4 uv4],
    "indices": u32, "sections": [(voff, vcount, ioff, icount, material)],
    "transform": 4x4}], "materials": [(base4, emissive3, metallic, roughness)],
    "camera_view": 4x4 or None}"""
    doc, binchunk = _load(path)
    bufs = _buffers(doc, path, binchunk)
    meshes = []
    for mi, mesh in enumerate(doc.get("meshes", [])):
        verts, inds, secs = [], [], []
        for p in mesh["primitives"]:
            at = p["attributes"]
            voff = sum(len(v) for v in verts)
            if "POSITION" in at:
                pos = _accessor(doc, bufs, at["POSITION"])
                n = len(pos)
                v = np.zeros((n, 16))
                v[:, 0:3] = pos
                v[:, 3] = 1.0
                v[:, 4:8] = 1.0
                v[:, 8:12] = [0.0, 1.0, 0.0, 1.0]
                if "NORMAL" in at:
                    v[:, 8:11] = _accessor(doc, bufs, at["NORMAL"])
                if "TEXCOORD_0" in at:
                    v[:, 12:14] = _accessor(doc, bufs, at["TEXCOORD_0"])
                if "COLOR_0" in at:
                    c = _accessor(doc, bufs, at["COLOR_0"])
                    v[:, 4:4 + c.shape[1]] = c
                verts.append(v)
            vcount = sum(len(v) for v in verts) - voff
            ioff = sum(len(i) for i in inds)
            icount = 0
            if "indices" in p:
                ii = _accessor(doc, bufs, p["indices"]).reshape(-1).astype(np.uint32)
                inds.append(ii)
                icount = len(ii)
            secs.append((voff, vcount, ioff, icount if "indices" in p else 0, p.get("material", -1)))
        meshes.append({
            "vertices": np.concatenate(verts) if verts else np.zeros((0, 16)),
            "indices": np.concatenate(inds) if inds else np.zeros(0, np.uint32),
            "sections": secs,
            "transform": _global_transform(doc, mi),
        })
    mats = []
    for m in doc.get("materials", []):
        pbr = m.get("pbrMetallicRoughness", {})
        mats.append((pbr.get("baseColorFactor", [1, 1, 1, 1]), m.get("emissiveFactor", [0, 0, 0]),
                     pbr.get("metallicFactor", 1.0), pbr.get("roughnessFactor", 1.0)))
    view = None
    cams = doc.get("cameras", [])
    if cams and cams[0].get("type") == "perspective":
        for n in doc.get("nodes", []):
            if n.get("camera") == 0:
                view = _node_matrix(n)
                break
    return {"meshes": meshes, "materials": mats, "camera_view": view}
