"""CPU tests of the glTF 2.0 asset reader (vkhrt_asset_*), the host side of ModelLoader
(source/resources/model/model_loader.cpp:139-312) and Node::GetWorldMatrix (model.cpp:6-18).

Files are written by tests/gltf_writer.py from known arrays, so every decoded value has an
exact expectation: positions, normals, texCoords and indices bit for bit; primitive topologies
against the glTF 2.0 specification's formulas; TRS transforms against scipy's quaternion
rotation in float64; world matrices bit for bit against the oracle's GetWorldMatrix and, on
dyadic matrices whose float products are exact, against float64 products."""
import json
import struct

import numpy as np
import pytest

import gltf_writer as GW

V = pytest.importorskip("vkhrt_amd")


def _load(g, how, tmp_path):
    if how == "bin":
        g.write_gltf(str(tmp_path / "a.gltf"))
        return V.Asset(str(tmp_path / "a.gltf"))
    if how == "data":
        return V.Asset(data=g.gltf_text(None).encode())
    p = tmp_path / "a.glb"
    p.write_bytes(g.glb())
    return V.Asset(str(p))


def _strands(rng, nstrands=40, nseg=6):
    v = np.zeros((nstrands, nseg + 1, 4), np.float32)
    v[..., :3] = np.cumsum(rng.normal(size=(nstrands, nseg + 1, 3)), axis=1).astype(np.float32)
    return GW.strand_lines_mesh(v, nstrands, nseg)


@pytest.mark.parametrize("how", ["bin", "data", "glb"])
def test_lines_round_trip(how, tmp_path):
    rng = np.random.default_rng(1)
    pos, idx = _strands(rng)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": pos, "indices": idx, "mode": 1}], name="hair"))])
    a = _load(g, how, tmp_path)
    info = a.info()
    assert info == {"mesh_count": 1, "node_count": 1, "material_count": 0, "instance_count": 1,
                    "line_mesh_count": 1, "triangle_mesh_count": 0}
    m = a.mesh(0)
    assert m["primitive"] == V.PRIMITIVE_LINES and m["material"] == -1 and m["name"] == "hair"
    assert np.array_equal(m["vertices"][:, :3].view(np.uint32), pos.view(np.uint32))
    assert not m["vertices"][:, 3:].any()  # no normals or texCoords in the file
    assert np.array_equal(m["indices"], idx)
    assert np.array_equal(m["bounds_min"], pos.min(0)) and np.array_equal(m["bounds_max"], pos.max(0))
    n = a.node(0)
    assert n["parent"] == -1 and n["meshes"] == [0]
    assert np.array_equal(n["world"], np.eye(4, dtype=np.float32).reshape(16))
    inst = a.instance(0)
    assert inst["node"] == 0 and inst["mesh"] == 0
    assert np.array_equal(inst["transform"], np.eye(4, dtype=np.float32)[:3].reshape(12))


def test_memory_load_resolves_external_buffers(tmp_path):
    rng = np.random.default_rng(2)
    pos, idx = _strands(rng, 5, 3)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": pos, "indices": idx, "mode": 1}]))])
    g.write_gltf(str(tmp_path / "x.gltf"), bin_name="my%20data.bin")
    (tmp_path / "my data.bin").write_bytes(bytes(g.blob))  # percent-decoded file name
    text = (tmp_path / "x.gltf").read_bytes()
    a = V.Asset(data=text, base_dir=str(tmp_path))
    assert np.array_equal(a.mesh(0)["indices"], idx)
    with pytest.raises(V.VkhrtError, match="no directory"):
        V.Asset(data=text)


@pytest.mark.parametrize("itype", [np.uint8, np.uint16, np.uint32])
@pytest.mark.parametrize("uv_type", [np.float32, np.uint8, np.uint16])
def test_interleaved_attributes_and_index_types(itype, uv_type):
    rng = np.random.default_rng(3)
    nv = 200
    pos = rng.normal(size=(nv, 3)).astype(np.float32)
    nrm = rng.normal(size=(nv, 3)).astype(np.float32)
    if uv_type == np.float32:
        uv = rng.random((nv, 2)).astype(np.float32)
        want_uv = uv
    else:
        hi = np.iinfo(uv_type).max
        uv = rng.integers(0, hi + 1, (nv, 2)).astype(uv_type)
        want_uv = uv.astype(np.float32) / np.float32(hi)  # the specification's c / max
    idx = rng.integers(0, nv, 99).astype(itype)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": pos, "normals": nrm, "uv": uv, "indices": idx, "interleave": True}]))])
    m = V.Asset(data=g.glb()).mesh(0)
    assert m["primitive"] == V.PRIMITIVE_TRIANGLES
    v = m["vertices"]
    assert np.array_equal(v[:, 0:3], pos) and np.array_equal(v[:, 3:6], nrm)
    assert np.array_equal(v[:, 6:8].view(np.uint32), want_uv.view(np.uint32))
    assert np.array_equal(m["indices"], idx.astype(np.uint32))


def _spec_indices(mode, seq):
    """glTF 2.0 specification, Meshes / primitive topology, written out per mode."""
    n = len(seq)
    out = []
    if mode == 1:
        for i in range(n // 2):
            out += [seq[2 * i], seq[2 * i + 1]]
    elif mode == 2:
        for i in range(n):
            out += [seq[i], seq[(i + 1) % n]]
    elif mode == 3:
        for i in range(n - 1):
            out += [seq[i], seq[i + 1]]
    elif mode == 4:
        for i in range(n // 3):
            out += seq[3 * i:3 * i + 3]
    elif mode == 5:
        for i in range(n - 2):
            out += [seq[i], seq[i + 1 + i % 2], seq[i + 2 - i % 2]]
    elif mode == 6:
        for i in range(n - 2):
            out += [seq[i + 1], seq[i + 2], seq[0]]
    return np.array(out, np.uint32)


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("indexed", [True, False])
def test_primitive_topologies(mode, indexed):
    rng = np.random.default_rng(10 + mode)
    nv = 13  # odd: LINES and TRIANGLES drop the trailing indices
    pos = rng.normal(size=(nv, 3)).astype(np.float32)
    seq = list(rng.permutation(nv)) if indexed else list(range(nv))
    prim = {"positions": pos, "mode": mode, "normals": np.zeros((nv, 3), np.float32)}
    if indexed:
        prim["indices"] = np.array(seq, np.uint16)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([prim]))])
    m = V.Asset(data=g.glb()).mesh(0)
    assert m["primitive"] == (V.PRIMITIVE_LINES if mode <= 3 else V.PRIMITIVE_TRIANGLES)
    assert np.array_equal(m["indices"], _spec_indices(mode, seq))


def test_generated_face_normals():
    """Triangle meshes without normals: per-face normals stored at the face's vertices, the
    last face listing a vertex winning; a flat quad gets its plane normal exactly."""
    quad = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0]], np.float32)
    g = GW.Gltf()
    mq = g.mesh([{"positions": quad, "indices": np.array([0, 1, 2, 0, 2, 3], np.uint16)}])
    rng = np.random.default_rng(4)
    tet = rng.normal(size=(4, 3)).astype(np.float32)
    faces = np.array([[0, 1, 2], [0, 3, 1], [1, 3, 2], [2, 3, 0]], np.uint32)
    mt = g.mesh([{"positions": tet, "indices": faces.reshape(-1)}])
    g.scene([g.node(mesh=mq), g.node(mesh=mt)])
    a = V.Asset(data=g.glb())
    assert np.array_equal(a.mesh(0)["vertices"][:, 3:6], np.tile([0, 0, 1], (4, 1)).astype(np.float32))
    want = np.zeros((4, 3))
    for f in faces:
        p = tet[f].astype(np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        want[f] = n / np.linalg.norm(n)
    assert np.abs(a.mesh(1)["vertices"][:, 3:6] - want).max() < 1e-6
    # lines keep zero normals (no normal is defined for them)
    g2 = GW.Gltf()
    g2.scene([g2.node(mesh=g2.mesh([{"positions": tet, "indices": np.array([0, 1, 2, 3], np.uint8), "mode": 1}]))])
    assert not V.Asset(data=g2.glb()).mesh(0)["vertices"][:, 3:].any()


def _rot(q):
    from scipy.spatial.transform import Rotation
    return Rotation.from_quat(q).as_matrix()


def _trs64(t, q, s):
    M = np.eye(4)
    M[:3, :3] = _rot(q) @ np.diag(s)
    M[:3, 3] = t
    return M


def _hierarchy(g, rng, dyadic=False):
    """root(TRS) -> a(matrix) -> b(TRS, mesh); root -> c(mesh). Returns the float64 world
    matrices per node name (row-major) and the local ones."""
    pos, idx = _strands(rng, 4, 3)
    mesh = g.mesh([{"positions": pos, "indices": idx, "mode": 1}])
    if dyadic:
        loc = {k: np.eye(4) for k in "rabc"}
        for k in loc:
            loc[k][:3, :3] = rng.integers(-2, 3, (3, 3)) * 0.5
            loc[k][:3, 3] = rng.integers(-8, 9, 3) * 0.25
        b = g.node(mesh=mesh, matrix=loc["b"].T, name="b")
        a = g.node(children=[b], matrix=loc["a"].T, name="a")
        c = g.node(mesh=mesh, matrix=loc["c"].T, name="c")
        r = g.node(children=[a, c], matrix=loc["r"].T, name="r")
    else:
        qs = {k: (lambda q: q / np.linalg.norm(q))(rng.normal(size=4)) for k in "rb"}
        loc = {"r": _trs64([1.0, -2.0, 0.5], qs["r"], [2.0, 2.0, 2.0]),
               "b": _trs64([0.0, 3.0, -1.0], qs["b"], [1.0, 0.5, 1.5]),
               "c": np.eye(4)}
        am = np.eye(4)
        am[:3, :3] = _rot(rng.normal(size=4) / 2.0)
        am[:3, 3] = [4.0, 0.0, 0.0]
        loc["a"] = am.astype(np.float32).astype(np.float64)
        b = g.node(mesh=mesh, translation=[0.0, 3.0, -1.0], rotation=qs["b"], scale=[1.0, 0.5, 1.5], name="b")
        a = g.node(children=[b], matrix=loc["a"].T, name="a")
        c = g.node(mesh=mesh, name="c")
        r = g.node(children=[a, c], translation=[1.0, -2.0, 0.5], rotation=qs["r"], scale=[2.0, 2.0, 2.0], name="r")
    g.scene([r])
    world = {"r": loc["r"], "a": loc["r"] @ loc["a"], "c": loc["r"] @ loc["c"]}
    world["b"] = world["a"] @ loc["b"]
    return world, loc


def test_node_hierarchy_and_world_matrices():
    rng = np.random.default_rng(5)
    g = GW.Gltf()
    world, loc = _hierarchy(g, rng)
    a = V.Asset(data=g.glb())
    nodes = a.nodes
    assert [n["name"] for n in nodes] == ["r", "a", "b", "c"]  # pre-order
    assert [n["parent"] for n in nodes] == [-1, 0, 1, 0]
    for n in nodes:
        L = n["local"].reshape(4, 4).T.astype(np.float64)  # stored column-major
        W = n["world"].reshape(4, 4).T.astype(np.float64)
        assert np.abs(L - loc[n["name"]]).max() < 2e-6 * max(1.0, np.abs(loc[n["name"]]).max())
        assert np.abs(W - world[n["name"]]).max() < 1e-5 * max(1.0, np.abs(world[n["name"]]).max())
    # bit for bit against the oracle's GetWorldMatrix over the loader's local matrices
    import oracle_lib as O
    ow = O.world_matrices(np.stack([n["local"] for n in nodes]), [n["parent"] for n in nodes])
    assert np.array_equal(ow.view(np.uint32), np.stack([n["world"] for n in nodes]).view(np.uint32))
    # the parent-first product order matters here: the reversed order is far from float64
    Wb_rev = loc["b"] @ loc["a"] @ loc["r"]
    assert np.abs(Wb_rev - world["b"]).max() > 1e-2
    # instances: (node, mesh) in pre-order, transform = rows 0..2 of the world matrix
    inst = a.instances
    assert [(i["node"], i["mesh"]) for i in inst] == [(2, 0), (3, 0)]
    for i in inst:
        W = nodes[i["node"]]["world"].reshape(4, 4).T
        assert np.array_equal(i["transform"], W[:3].reshape(12))


def test_world_matrices_exact_on_dyadic_transforms():
    """Small dyadic entries make every float product exact: the world matrices then equal the
    float64 products bit for bit, in the loader and in the oracle."""
    rng = np.random.default_rng(6)
    g = GW.Gltf()
    world, _ = _hierarchy(g, rng, dyadic=True)
    nodes = V.Asset(data=g.glb()).nodes
    for n in nodes:
        assert np.array_equal(n["world"].reshape(4, 4).T, world[n["name"]].astype(np.float32))
    import oracle_lib as O
    ow = O.world_matrices(np.stack([n["local"] for n in nodes]), [n["parent"] for n in nodes])
    for n, w in zip(nodes, ow):
        assert np.array_equal(w.reshape(4, 4).T, world[n["name"]].astype(np.float32))
    with pytest.raises(ValueError):
        O.world_matrices(np.zeros((2, 16), np.float32), [1, 0])


def test_roots_scenes_and_multi_primitive_meshes():
    rng = np.random.default_rng(7)
    pos, idx = _strands(rng, 3, 2)
    g = GW.Gltf()
    m0 = g.mesh([{"positions": pos, "indices": idx, "mode": 1},
                 {"positions": pos[:5], "indices": np.array([0, 1, 1, 2, 3, 4], np.uint8), "mode": 1}])
    m1 = g.mesh([{"positions": pos, "indices": idx, "mode": 3}])
    n0 = g.node(mesh=m0, name="n0", translation=[1, 0, 0])
    n1 = g.node(mesh=m1, name="n1")
    n2 = g.node(mesh=m0, name="n2", scale=[2, 2, 2])
    g.scene([n0, n1], default=False)
    g.scene([n2, n0])
    a = V.Asset(data=g.glb())
    info = a.info()
    assert info["mesh_count"] == 3 and info["line_mesh_count"] == 3
    nodes = a.nodes
    # two scene roots: an identity "ROOT" holds them; the default scene (1) is the one read
    assert [n["name"] for n in nodes] == ["ROOT", "n2", "n0"]
    assert [n["parent"] for n in nodes] == [-1, 0, 0]
    assert nodes[1]["meshes"] == [0, 1] and nodes[2]["meshes"] == [0, 1]
    assert [(i["node"], i["mesh"]) for i in a.instances] == [(1, 0), (1, 1), (2, 0), (2, 1)]
    assert a.instance(2)["transform"][3] == 1.0 and a.instance(0)["transform"][0] == 2.0
    # no scenes: every parentless node is a root
    g.doc["scenes"] = []
    g.doc.pop("scene")
    assert [n["name"] for n in V.Asset(data=g.glb()).nodes] == ["ROOT", "n0", "n1", "n2"]


def test_materials():
    g = GW.Gltf()
    m0 = g.material(base_color=[0.5, 0.25, 1.0, 1.0], metallic=0.0, roughness=0.75, emissive=[1.0, 0.0, 0.5])
    g.material()
    pos = np.eye(3, dtype=np.float32)
    mesh = g.mesh([{"positions": pos, "material": m0}, {"positions": pos}, {"positions": pos, "material": 1}])
    g.scene([g.node(mesh=mesh)])
    a = V.Asset(data=g.glb())
    assert [a.mesh(i)["material"] for i in range(3)] == [0, -1, 1]
    assert a.material(0) == {"base_color": [0.5, 0.25, 1.0, 1.0], "metallic": 0.0, "roughness": 0.75,
                             "emissive": [1.0, 0.0, 0.5]}
    assert a.material(1) == {"base_color": [1.0] * 4, "metallic": 1.0, "roughness": 1.0, "emissive": [0.0] * 3}


def test_unicode_names_and_json_escapes():
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": np.eye(3, dtype=np.float32)}]), name="très \U0001F600 \"q\"")])
    text = g.gltf_text(None)
    assert "\\ud83d\\ude00" in text  # json.dumps escapes non-ASCII, surrogate pairs included
    assert V.Asset(data=text.encode()).node(0)["name"] == "très \U0001F600 \"q\""


def _doc_with(mutate):
    rng = np.random.default_rng(8)
    pos, idx = _strands(rng, 3, 2)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": pos, "indices": idx, "mode": 1}]))])
    d = json.loads(g.gltf_text(None))
    mutate(d)
    return json.dumps(d).encode()


def _set(path, value):
    def f(d):
        o = d
        for k in path[:-1]:
            o = o[k]
        o[path[-1]] = value
    return f


ERRORS = [
    ("version", _set(["asset", "version"], "1.0"), "UNSUPPORTED", "version"),
    ("index", lambda d: d["accessors"].append({"bufferView": 0, "componentType": 5126, "count": 1, "type": "SCALAR"})
     or _set(["meshes", 0, "primitives", 0, "indices"], 0)(d), "INVALID_ARGUMENT", "unsigned"),
    ("overrun", _set(["accessors", 0, "count"], 10 ** 6), "INVALID_ARGUMENT", "overruns"),
    ("sparse", _set(["accessors", 0, "sparse"], {"count": 1}), "UNSUPPORTED", "sparse"),
    ("points", _set(["meshes", 0, "primitives", 0, "mode"], 0), "UNSUPPORTED", "POINTS"),
    ("no position", _set(["meshes", 0, "primitives", 0, "attributes"], {}), "INVALID_ARGUMENT", "POSITION"),
    ("two parents", lambda d: d["nodes"].extend([{"children": [0]}, {"children": [0]}]), "INVALID_ARGUMENT", "parent"),
    ("bad base64", _set(["buffers", 0, "uri"], "data:application/octet-stream;base64,@@@@"), "INVALID_ARGUMENT", "base64"),
    ("short buffer", _set(["buffers", 0, "byteLength"], 10 ** 7), "INVALID_ARGUMENT", "byteLength"),
    ("material", _set(["meshes", 0, "primitives", 0, "material"], 3), "INVALID_ARGUMENT", "material"),
    ("mesh ref", _set(["nodes", 0, "mesh"], 9), "INVALID_ARGUMENT", "mesh index"),
]


@pytest.mark.parametrize("name,mutate,code,msg", ERRORS, ids=[e[0] for e in ERRORS])
def test_malformed_assets_are_rejected(name, mutate, code, msg):
    data = _doc_with(mutate)
    with pytest.raises(V.VkhrtError) as e:
        V.Asset(data=data)
    assert V.STATUS[e.value.code] == code
    assert msg in str(e.value)


def test_out_of_range_index_is_rejected():
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": np.eye(3, dtype=np.float32),
                                  "indices": np.array([0, 1, 3, 2], np.uint16), "mode": 1}]))])
    with pytest.raises(V.VkhrtError, match="not below the vertex count 3"):
        V.Asset(data=g.glb())


@pytest.mark.parametrize("data", [b"", b"{", b"[1, 2", b'{"asset": {"version": "2.0"}} x', b"\xff\xfe",
                                  b"[" * 300 + b"]" * 300, b'{"a": "\\ud800"}',
                                  struct.pack("<III", 0x46546C67, 2, 1000),
                                  struct.pack("<IIIII", 0x46546C67, 2, 20, 100, 0x4E4F534A)])
def test_garbage_is_rejected_without_crashing(data):
    with pytest.raises(V.VkhrtError):
        V.Asset(data=data)


def test_missing_file(tmp_path):
    with pytest.raises(V.VkhrtError, match="cannot read"):
        V.Asset(str(tmp_path / "nope.gltf"))
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": np.eye(3, dtype=np.float32)}]))])
    (tmp_path / "a.gltf").write_text(g.gltf_text("gone.bin"))
    with pytest.raises(V.VkhrtError, match="gone.bin"):
        V.Asset(str(tmp_path / "a.gltf"))


def test_empty_scene_has_an_identity_root():
    d = b'{"asset": {"version": "2.0"}}'
    a = V.Asset(data=d)
    assert a.info()["node_count"] == 1 and a.info()["instance_count"] == 0
    assert a.node(0)["name"] == "ROOT"


def test_mesh_lines_oracle_and_the_reference_offset():
    """GenerateLines over mesh-local indices (the oracle's restatement) is a gather of the
    mesh's own positions. With two meshes in one vertex buffer the reference adds the second
    mesh's firstVertex twice (model_loader.cpp:157 and geometry_processor.cpp:61), reading
    other vertices; the loader keeps indices mesh-local."""
    import oracle_lib as O
    rng = np.random.default_rng(9)
    p0, i0 = _strands(rng, 6, 4)
    p1, i1 = _strands(rng, 5, 3)
    g = GW.Gltf()
    g.scene([g.node(mesh=g.mesh([{"positions": p0, "indices": i0, "mode": 1},
                                 {"positions": p1, "indices": i1, "mode": 1}]))])
    a = V.Asset(data=g.glb())
    m1 = a.mesh(1)
    lines = O.mesh_lines(m1["vertices"], m1["indices"], radius=0.25)
    assert np.array_equal(lines[:, 0:3], p1[i1[0::2]]) and np.array_equal(lines[:, 4:7], p1[i1[1::2]])
    assert np.all(lines[:, 3] == 0.25) and np.all(lines[:, 7] == 0.25)
    # the reference's concatenated buffers and doubly offset indices
    allpos = np.concatenate([p0, p1])
    first = p0.shape[0]
    stored = i1 + first
    twice = stored + first
    assert twice.max() >= allpos.shape[0] or not np.array_equal(allpos[twice[0::2]], p1[i1[0::2]])
    with pytest.raises(ValueError):
        O.mesh_lines(m1["vertices"], np.array([0, m1["vertices"].shape[0]], np.uint32))


def test_cpp_model_loader_mirror(tmp_path):
    """include/vkhrt.hpp ModelLoader / Model / SceneGraph (host only) report the same scene graph
    as the C-ABI: meshes, index arrays, bounds, node parents and world matrices bit for bit."""
    import subprocess
    from test_host import cpp_test_binary
    rng = np.random.default_rng(11)
    g = GW.Gltf()
    _hierarchy(g, rng)
    g.material(base_color=[1.0, 0.5, 0.5, 1.0])
    p = tmp_path / "h.glb"
    p.write_bytes(g.glb())
    r = subprocess.run([cpp_test_binary(), "assetinfo", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "ok" and lines[-2] == "materials 1"
    a = V.Asset(str(p))
    meshes = [ln.split() for ln in lines if ln.startswith("mesh ")]
    nodes = [ln.split() for ln in lines if ln.startswith("node ")]
    assert len(meshes) == a.info()["mesh_count"] and len(nodes) == a.info()["node_count"]
    for fields, i in zip(meshes, range(len(meshes))):
        m = a.mesh(i)
        digest = 0
        for x in m["indices"].tolist():
            digest = (digest * 1000003 + x) % (1 << 64)
        assert fields[1:6] == [str(m["primitive"]), str(m["vertices"].shape[0]), str(m["indices"].size),
                               str(m["material"]), str(digest)]
        assert float.fromhex(fields[6]) == m["bounds_min"][0] and float.fromhex(fields[7]) == m["bounds_max"][2]
    for fields, n in zip(nodes, a.nodes):
        assert int(fields[1]) == n["parent"] and int(fields[2]) == len(n["meshes"])
        w = np.array([float.fromhex(x) for x in fields[3:]], np.float32)
        assert np.array_equal(w.view(np.uint32), n["world"].view(np.uint32))
    r = subprocess.run([cpp_test_binary(), "assetinfo", str(tmp_path / "none.glb")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("error -1 ") and "cannot read" in r.stdout


def _fuzz_cases(seed, n):
    """Valid files (a GLB and a .gltf with a data: URI, both with nodes, materials and strided
    attributes) with random byte edits: flips, overwrites with JSON punctuation or digits,
    truncations and duplications."""
    rng = np.random.default_rng(seed)
    pos, idx = _strands(rng, 6, 3)
    g = GW.Gltf()
    m = g.material(base_color=[0.5, 0.5, 0.5, 1.0])
    mesh = g.mesh([{"positions": pos, "indices": idx, "mode": 1, "material": m},
                   {"positions": pos[:9], "normals": pos[:9], "uv": np.zeros((9, 2), np.float32),
                    "indices": np.arange(9, dtype=np.uint16), "interleave": True}])
    child = g.node(mesh=mesh, translation=[1.0, 2.0, 3.0], rotation=[0.0, 0.0, 0.6, 0.8])
    g.scene([g.node(children=[child], scale=[2.0, 2.0, 2.0]), g.node(mesh=mesh)])
    bases = [bytearray(g.glb()), bytearray(g.gltf_text(None).encode())]
    punct = b'{}[]":,0123456789-.eE '
    for _ in range(n):
        b = bytearray(bases[int(rng.integers(2))])
        for _ in range(int(rng.integers(1, 6))):
            k = int(rng.integers(5))
            i = int(rng.integers(len(b)))
            if k == 0:
                b[i] ^= 1 << int(rng.integers(8))
            elif k == 1:
                b[i] = punct[int(rng.integers(len(punct)))]
            elif k == 2:
                b = b[:i]
            elif k == 3:
                j = min(len(b), i + int(rng.integers(1, 16)))
                b[i:i] = b[i:j]
            else:
                b[i:i + 4] = int(rng.integers(0, 2 ** 32)).to_bytes(4, "little")
            if not b:
                break
        yield bytes(b)


def test_mutated_files_load_or_fail_cleanly():
    """Every mutated file either loads into a consistent asset or fails with a status and a
    message; the reader never crashes or reads outside the data (a crash ends the test run)."""
    loaded = failed = 0
    for data in _fuzz_cases(17, 3000):
        try:
            a = V.Asset(data=data)
        except V.VkhrtError as e:
            assert e.code in (-1, -4) and str(e)
            failed += 1
            continue
        info = a.info()
        for i in range(info["mesh_count"]):
            m = a.mesh(i)
            nv = m["vertices"].shape[0]
            assert m["indices"].size == 0 or int(m["indices"].max()) < nv
            assert m["indices"].size % (2 if m["primitive"] == V.PRIMITIVE_LINES else 3) == 0
        for i in range(info["node_count"]):
            n = a.node(i)
            assert -1 <= n["parent"] < i
            assert all(0 <= k < info["mesh_count"] for k in n["meshes"])
        assert info["instance_count"] == sum(len(a.node(i)["meshes"]) for i in range(info["node_count"]))
        loaded += 1
    print(f"fuzz: {loaded} loaded, {failed} rejected")
    assert loaded > 100 and failed > 100


def test_mutated_files_under_address_sanitizer(tmp_path):
    """The same mutated files through the reader built with AddressSanitizer and
    UndefinedBehaviorSanitizer (tests/cpp/asset_fuzz, host code only): no invalid access."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "cpp", "asset_fuzz")
    subprocess.run(["make", "-s", "-C", os.path.join(here, "cpp"), exe], check=True)
    paths = []
    for k, data in enumerate(_fuzz_cases(23, 1500)):
        p = tmp_path / f"m{k}.glb"
        p.write_bytes(data)
        paths.append(str(p))
    r = subprocess.run([exe] + paths, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1"))
    assert r.returncode == 0, r.stderr[-3000:]
    loaded, rejected = int(r.stdout.split()[0]), int(r.stdout.split()[2])
    assert loaded + rejected == len(paths) and loaded > 50
