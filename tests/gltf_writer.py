"""A small glTF 2.0 encoder for the asset tests: meshes, nodes, materials and scenes from numpy
arrays, written as .gltf + .bin, as .gltf with base64 data: URIs, or as .glb. It follows the
layout rules of the glTF 2.0 specification and shares no code with the library's reader, so a
round trip through the reader checks the reader's decoding."""
import base64
import json
import os
import struct

import numpy as np

COMPONENT = {np.dtype(np.int8): 5120, np.dtype(np.uint8): 5121, np.dtype(np.int16): 5122,
             np.dtype(np.uint16): 5123, np.dtype(np.uint32): 5125, np.dtype(np.float32): 5126}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}


class Gltf:
    def __init__(self):
        self.doc = {"asset": {"version": "2.0", "generator": "vkhrt tests"}, "buffers": [], "bufferViews": [],
                    "accessors": [], "meshes": [], "nodes": [], "materials": [], "scenes": []}
        self.blob = bytearray()

    def _align(self, n=4):
        while len(self.blob) % n:
            self.blob += b"\0"

    def view(self, data, stride=None):
        self._align()
        v = {"buffer": 0, "byteOffset": len(self.blob), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.blob += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, normalized=False, view=None, offset=0):
        a = np.ascontiguousarray(arr)
        k = 1 if a.ndim == 1 else a.shape[1]
        if view is None:
            view = self.view(a.tobytes())
        acc = {"bufferView": view, "componentType": COMPONENT[a.dtype], "count": int(a.shape[0]), "type": TYPE[k]}
        if offset:
            acc["byteOffset"] = offset
        if normalized:
            acc["normalized"] = True
        if a.dtype == np.float32 and k == 3 and a.shape[0]:
            acc["min"] = [float(x) for x in a.min(0)]
            acc["max"] = [float(x) for x in a.max(0)]
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def interleaved(self, columns):
        """One strided buffer view holding every column's element in turn; returns accessors."""
        n = columns[0].shape[0]
        cols = [np.ascontiguousarray(c) for c in columns]
        sizes = [c.dtype.itemsize * (1 if c.ndim == 1 else c.shape[1]) for c in cols]
        offs = np.cumsum([0] + [(s + 3) // 4 * 4 for s in sizes])
        stride = int(offs[-1])
        raw = np.zeros((n, stride), np.uint8)
        for c, s, o in zip(cols, sizes, offs):
            raw[:, o:o + s] = c.view(np.uint8).reshape(n, s)
        v = self.view(raw.tobytes(), stride=stride)
        return [self.accessor(c, view=v, offset=int(o)) for c, o in zip(cols, offs)]

    def mesh(self, primitives, name=None):
        """primitives: dicts with positions (n, 3) float32 and optional normals, uv, indices,
        mode (glTF default 4 when absent), material, interleave (all attributes in one strided
        view). Integer uv arrays are written as normalized."""
        prims = []
        for p in primitives:
            attrs = {}
            cols = [("POSITION", np.asarray(p["positions"], np.float32))]
            if p.get("normals") is not None:
                cols.append(("NORMAL", np.asarray(p["normals"], np.float32)))
            if p.get("uv") is not None:
                cols.append(("TEXCOORD_0", np.asarray(p["uv"])))
            if p.get("interleave"):
                accs = self.interleaved([c for _, c in cols])
                for (key, c), a in zip(cols, accs):
                    attrs[key] = a
                    if key == "TEXCOORD_0" and c.dtype != np.float32:
                        self.doc["accessors"][a]["normalized"] = True
            else:
                for key, c in cols:
                    attrs[key] = self.accessor(c, normalized=key == "TEXCOORD_0" and c.dtype != np.float32)
            prim = {"attributes": attrs}
            if p.get("indices") is not None:
                prim["indices"] = self.accessor(np.asarray(p["indices"]))
            if p.get("mode") is not None:
                prim["mode"] = int(p["mode"])
            if p.get("material") is not None:
                prim["material"] = int(p["material"])
            prims.append(prim)
        m = {"primitives": prims}
        if name:
            m["name"] = name
        self.doc["meshes"].append(m)
        return len(self.doc["meshes"]) - 1

    def node(self, mesh=None, children=None, matrix=None, translation=None, rotation=None, scale=None, name=None):
        n = {}
        if mesh is not None:
            n["mesh"] = mesh
        if children:
            n["children"] = list(children)
        if matrix is not None:
            n["matrix"] = [float(x) for x in np.asarray(matrix, np.float64).reshape(16)]
        for key, val in (("translation", translation), ("rotation", rotation), ("scale", scale)):
            if val is not None:
                n[key] = [float(x) for x in val]
        if name is not None:
            n["name"] = name
        self.doc["nodes"].append(n)
        return len(self.doc["nodes"]) - 1

    def material(self, base_color=None, metallic=None, roughness=None, emissive=None):
        m, pbr = {}, {}
        if base_color is not None:
            pbr["baseColorFactor"] = list(base_color)
        if metallic is not None:
            pbr["metallicFactor"] = metallic
        if roughness is not None:
            pbr["roughnessFactor"] = roughness
        if pbr:
            m["pbrMetallicRoughness"] = pbr
        if emissive is not None:
            m["emissiveFactor"] = list(emissive)
        self.doc["materials"].append(m)
        return len(self.doc["materials"]) - 1

    def scene(self, roots, default=True):
        self.doc["scenes"].append({"nodes": list(roots)})
        if default:
            self.doc["scene"] = len(self.doc["scenes"]) - 1
        return len(self.doc["scenes"]) - 1

    def _doc(self, uri):
        d = json.loads(json.dumps(self.doc))
        for key in ("materials", "scenes", "meshes", "nodes", "accessors", "bufferViews"):
            if not d[key]:
                del d[key]
        d["buffers"] = [{"byteLength": len(self.blob)}] if self.blob else []
        if uri is not None and self.blob:
            d["buffers"][0]["uri"] = uri
        if not d["buffers"]:
            del d["buffers"]
        return d

    def gltf_text(self, bin_name=None):
        """.gltf JSON; buffer as external file `bin_name`, or base64 data: URI if None."""
        uri = bin_name if bin_name else "data:application/octet-stream;base64," + base64.b64encode(
            bytes(self.blob)).decode()
        return json.dumps(self._doc(uri), indent=1)

    def write_gltf(self, path, bin_name="buffer.bin"):
        with open(path, "w") as f:
            f.write(self.gltf_text(bin_name))
        if bin_name:
            with open(os.path.join(os.path.dirname(path), bin_name), "wb") as f:
                f.write(bytes(self.blob))

    def glb(self):
        js = json.dumps(self._doc(None)).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        bn = bytes(self.blob) + b"\0" * ((4 - len(self.blob) % 4) % 4)
        chunks = struct.pack("<II", len(js), 0x4E4F534A) + js
        if self.blob:
            chunks += struct.pack("<II", len(bn), 0x004E4942) + bn
        return struct.pack("<III", 0x46546C67, 2, 12 + len(chunks)) + chunks


def strand_lines_mesh(verts, nstrands, nseg):
    """Polyline strands (nstrands x (nseg + 1) vertices) as a LINES primitive: positions and
    the index pairs (k, k + 1) within each strand."""
    pos = np.ascontiguousarray(verts.reshape(-1, 4)[:, :3], np.float32)
    base = (np.arange(nstrands) * (nseg + 1))[:, None]
    k = np.arange(nseg)[None, :]
    idx = np.stack([base + k, base + k + 1], axis=-1).reshape(-1).astype(np.uint32)
    return pos, idx
