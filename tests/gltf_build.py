"""Test-side glTF 2.0 writer: builds the JSON document and one binary buffer,
and writes it as .glb, as .gltf with an external .bin, or as .gltf with a
data: URI, so the loader's readers can be checked against each other."""
import base64
import json
import struct

import numpy as np

COMPONENT = {np.dtype("i1"): 5120, np.dtype("u1"): 5121, np.dtype("i2"): 5122, np.dtype("u2"): 5123,
             np.dtype("u4"): 5125, np.dtype("f4"): 5126}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}


class GltfBuilder:
    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [],
                    "meshes": [], "nodes": [], "scenes": [{"nodes": []}], "materials": [],
                    "textures": [], "images": []}

    def _align(self):
        while len(self.bin) % 4:
            self.bin.append(0)

    def view(self, data, stride=None):
        """Appends raw bytes as a buffer view; returns its index."""
        self._align()
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, view=None, offset=0, count=None, width=None, dtype=None):
        """An accessor over `arr` (its own tight view), or over an existing
        view at byte `offset` with `count` elements of `width` x `dtype`."""
        if view is None:
            a = np.ascontiguousarray(arr)
            width = 1 if a.ndim == 1 else a.shape[1]
            dtype, count = a.dtype, len(a)
            view = self.view(a.tobytes())
        acc = {"bufferView": view, "byteOffset": offset, "componentType": COMPONENT[np.dtype(dtype)],
               "count": count, "type": TYPE[width]}
        if arr is not None and np.dtype(dtype) == np.float32 and width == 3:
            a = np.asarray(arr, np.float32).reshape(-1, 3)
            acc["min"], acc["max"] = a.min(0).tolist(), a.max(0).tolist()
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def primitive(self, positions, indices=None, normals=None, uvs=None, material=None, mode=None, index_dtype="u2"):
        at = {"POSITION": self.accessor(np.asarray(positions, np.float32))}
        if normals is not None:
            at["NORMAL"] = self.accessor(np.asarray(normals, np.float32))
        if uvs is not None:
            at["TEXCOORD_0"] = self.accessor(np.asarray(uvs, np.float32))
        p = {"attributes": at}
        if indices is not None:
            p["indices"] = self.accessor(np.asarray(indices, index_dtype))
        if material is not None:
            p["material"] = material
        if mode is not None:
            p["mode"] = mode
        return p

    def mesh(self, primitives):
        self.doc["meshes"].append({"primitives": primitives})
        return len(self.doc["meshes"]) - 1

    def node(self, root=False, **kw):
        self.doc["nodes"].append(kw)
        i = len(self.doc["nodes"]) - 1
        if root:
            self.doc["scenes"][0]["nodes"].append(i)
        return i

    def material(self, **kw):
        self.doc["materials"].append(kw)
        return len(self.doc["materials"]) - 1

    def image_png(self, png, as_uri=False):
        if as_uri:
            self.doc["images"].append({"uri": "data:image/png;base64," + base64.b64encode(png).decode()})
        else:
            self.doc["images"].append({"bufferView": self.view(png), "mimeType": "image/png"})
        self.doc["textures"].append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def _json(self, buffer):
        d = dict(self.doc)
        d["buffers"] = [buffer]
        for k in ("materials", "textures", "images"):
            if not d[k]:
                del d[k]
        return d

    def write_glb(self, path):
        self._align()
        js = json.dumps(self._json({"byteLength": len(self.bin)})).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(self.bin), 0x004E4942) + bytes(self.bin)
        path.write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
        return path

    def write_gltf(self, path, data_uri=False):
        self._align()
        if data_uri:
            buf = {"byteLength": len(self.bin), "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()}
        else:
            name = path.stem + ".bin"
            (path.parent / name).write_bytes(bytes(self.bin))
            buf = {"byteLength": len(self.bin), "uri": name}
        path.write_text(json.dumps(self._json(buf)))
        return path
