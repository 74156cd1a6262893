"""Expectations of the reference's light transport that the mathematics fixes
without the renderer's code, and the scenes they are taken on.

* direct light from a rectangular emitter onto an opaque plane: the
  direct-lighting integral by Gauss-Legendre quadrature over the emitter
  (next-event estimation, pathtrace.metal:319-384; with MIS the reference's
  light-hit weight, :589-637, as a four-dimensional integral); the quadrature
  itself is checked against Lambert's closed-form polygon form factor;
* a furnace: a closed sphere whose walls all emit E, seen from inside; by the
  chord symmetry of the sphere the radiance depends only on the angle to the
  wall, so the expectation of a depth-D estimate is E * sum_{k<=D} (T^k 1)(mu)
  with T the one-dimensional transport operator of the BSDF (the bounce loop
  with roulette, :552-728);
* a smooth glass slab between two emitting walls: the Fresnel chain of
  reflections and transmissions (:470-534) gives E_top F + E_bottom (1 - F)^2
  at depth 3, with F the exact unpolarised Fresnel reflectance;
* the hue shader (:193-201): the colour of hsv2rgb(h, 0.8, 1) from Python's
  colorsys;
* the denoiser (:1263-1354) and the ReSTIR merge (:919-1054) on constructed
  G-buffers whose answers are box filters, single edge-stopping weights,
  clamped blends and sums of reservoir weights.

The opaque BSDF the integrals use is the reference's microfacet model in its
textbook form (Lambert weighted by 1 - Fresnel, GGX distribution, separable
Smith shadowing, Schlick's Fresnel; :243-295): the checks pin how the
estimators sample, weight and accumulate it, and the BSDF itself is pinned
separately (reciprocity, pdf mass, Fresnel closed forms in
test_oracle_pins.py).
"""
from __future__ import annotations

import colorsys
import math

import numpy as np

import raymotion_amd as rm
from helpers import frame
from raymotion_amd import rng as np_rng
from raymotion_amd import scenes
from raymotion_amd.scene import material_rows

LUM = np.array([0.2126, 0.7152, 0.0722])


# ---------------------------------------------------------------- cameras
def look_at(pos, target, focal, up=(0.0, 0.0, 1.0)) -> rm.Camera:
    """Camera at pos looking at target; sensor x along right, y along up."""
    pos = np.asarray(pos, np.float64)
    fwd = np.asarray(target, np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    m = np.stack([right, upv, fwd], axis=1)  # columns: right, up, forward
    c = rm.Camera()
    c.x, c.y, c.z = (float(v) for v in pos)
    c.focal = float(focal)
    c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, c.m6, c.m7, c.m8 = (float(v) for v in m.reshape(9))
    return c


def corner_rays(cam: rm.Camera, width: int, height: int):
    """Primary rays of a frame with resolution 0: every sample of pixel (x, y)
    leaves through the sensor point (x - W/2, H/2 - y) (pathtrace.metal:751-752)."""
    y, x = np.mgrid[0:height, 0:width]
    sx = (x - width * 0.5).reshape(-1).astype(np.float64)
    sy = (height * 0.5 - y).reshape(-1).astype(np.float64)
    t = np.stack([sx, sy, np.full_like(sx, cam.focal)], axis=1)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    m = np.array([cam.m0, cam.m1, cam.m2, cam.m3, cam.m4, cam.m5, cam.m6, cam.m7, cam.m8], np.float64).reshape(3, 3)
    d = t @ m.T
    o = np.broadcast_to(np.array([cam.x, cam.y, cam.z], np.float64), d.shape)
    return o, d


def pin_frame(width, height, spp, bounces, cam, *, nomis=1, candidates=1, restir=False):
    """Frame settings for the transport pins: corner sampling (resolution 0),
    no clamp, no adaptive stop; ReSTIR path with its spatial pass reduced to
    the identity (radius 0.1 px: every neighbour rounds to the pixel itself),
    no temporal reuse and no denoiser when `restir`."""
    f = frame(width, height, spp, bounces, cam=cam, candidates=candidates, spatial=1 if restir else 0,
              radius=0.1, clamp=0.0, adapt_min=0, denoise=0, temporal=0, nomis=nomis)
    f.resolution = 0.0
    return f


# ---------------------------------------------------------------- BSDF
def spec_prob(metallic):
    sp = min(0.05 + 0.95 * metallic, 0.98)
    return 1.0 if metallic >= 0.999 else sp


def opaque_bsdf(wi, wo, n, rough, metallic, albedo):
    """Microfacet BSDF value (..., 3) and sampling pdf (...) for unit wi, wo, n."""
    rough = min(max(rough, 0.02), 1.0)
    a2 = rough ** 4
    albedo = np.asarray(albedo, np.float64)
    ci = np.sum(wi * n, -1)
    co = np.sum(wo * n, -1)
    h = wi + wo
    h = h / np.maximum(np.linalg.norm(h, axis=-1, keepdims=True), 1e-300)
    ch = np.abs(np.sum(h * n, -1))
    voh = np.abs(np.sum(wo * h, -1))
    d = a2 / (np.pi * (ch * ch * (a2 - 1.0) + 1.0) ** 2)

    def g1(c):
        return 2.0 * c / (c + np.sqrt(c * c + a2 * (1.0 - c * c)))

    f0 = 0.04 * (1.0 - metallic) + albedo * metallic
    fr = f0 + (1.0 - f0) * ((1.0 - voh) ** 5)[..., None]
    kd = albedo * (1.0 - metallic) * (1.0 - fr)
    with np.errstate(divide="ignore", invalid="ignore"):
        spec = (d * g1(ci) * g1(co) / (4.0 * ci * co))[..., None] * fr
        val = kd / np.pi + spec
        sp = spec_prob(metallic)
        pdf = (1.0 - sp) * ci / np.pi + sp * d * ch / (4.0 * voh)
    ok = (ci > 1e-8) & (co > 1e-8)
    val = np.where(ok[..., None], val, 0.0)
    pdf = np.where(ok, pdf, 0.0)
    return val, pdf


# ---------------------------------------------------------------- quadrature
def gauss_square(n):
    """Gauss-Legendre nodes (s, t) on [0, 1]^2 and their weights."""
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    s, t = np.meshgrid(x, x, indexing="ij")
    return s.reshape(-1), t.reshape(-1), np.outer(w, w).reshape(-1)


def polygon_form_factor_integral(x, n, verts):
    """Lambert's closed form of the integral over a planar polygon of
    cos(at x) cos(at the polygon) / r^2 dA, seen from point x with normal n:
    (1/2) sum_i theta_i (n . unit(R_i x R_{i+1})), R_i = v_i - x; the polygon
    wound so that its normal faces x."""
    total = 0.0
    k = len(verts)
    for i in range(k):
        a = np.asarray(verts[i], np.float64) - x
        b = np.asarray(verts[(i + 1) % k], np.float64) - x
        cr = np.cross(a, b)
        theta = math.atan2(np.linalg.norm(cr), float(np.dot(a, b)))
        total += theta * float(np.dot(n, cr / np.linalg.norm(cr)))
    return 0.5 * abs(total)


# ---------------------------------------------------------------- direct light scene
class DirectLight:
    """An opaque plane z = 0 lit by a rectangular emitter facing down; the
    camera sits below the emitter and looks down at the plane."""
    albedo = (0.8, 0.5, 0.3)
    metallic = 0.3
    rough = 0.6
    emission = np.array([4.0, 3.0, 2.0])
    corner = np.array([0.0, -0.1, 1.0])
    e1 = np.array([0.6, 0.0, 0.0])
    e2 = np.array([0.0, 0.4, 0.0])

    def __init__(self):
        c, e1, e2 = self.corner, self.e1, self.e2
        s = rm.Scene()
        plane = scenes._quad((-6, -6, 0), (6, -6, 0), (6, 6, 0), (-6, 6, 0))
        up = np.array([[0.0, 0.0, 1.0]] * 3)
        s.add_triangles(plane, material_rows(2, albedo=self.albedo, metallic=self.metallic, rough=self.rough),
                        normals=np.array([up, up]))
        # wound so that (v1 - v0) x (v2 - v0) points down, towards the plane
        light = scenes._quad(c, c + e2, c + e1 + e2, c + e1)
        down = np.array([[0.0, 0.0, -1.0]] * 3)
        s.add_triangles(light, material_rows(2, albedo=(0, 0, 0), emission=tuple(self.emission)),
                        normals=np.array([down, down]))
        s.build_bvh()
        self.exported = s.export()

    @staticmethod
    def camera(width):
        # below the emitter (z 0.6 < 1), looking down at 35 degrees with a
        # 2 x 20 degree field of view: the emitter stays out of view
        return look_at((-0.9, -0.9, 0.6), (-0.3, -0.3, 0.0), (width / 2) / math.tan(math.radians(20.0)))

    def hits(self, cam, width, height):
        o, d = corner_rays(cam, width, height)
        t = -o[:, 2] / d[:, 2]
        x = o + t[:, None] * d
        assert np.all(t > 0) and np.all(np.abs(x[:, :2]) < 5.0), "pixels must see the plane"
        return x, -d

    def light_points(self, n):
        s, t, w = gauss_square(n)
        area = float(np.linalg.norm(np.cross(self.e1, self.e2)))
        y = self.corner + s[:, None] * self.e1 + t[:, None] * self.e2
        return y, w * area, area

    def visible_points(self, xp, n=32):
        """Quadrature nodes and weights (area included) over the part of the
        emitter that next-event estimation sees from plane point xp.

        The reference offsets the shadow ray's origin by 0.002 along the
        shading normal but keeps the direction and length (dist - 0.003) of
        the unshifted segment (pathtrace.metal:348-351). From a point below a
        parallel emitter at height H the shifted ray meets the emitter's plane
        at (H - 0.002) / d_z, which is before its end when d_z < 2/3; it is
        then blocked wherever the crossing point y - (0.002 / H)(y - xp)
        (horizontally) lies on the emitter. So a sample y is seen when
        |y - xp|_horizontal <= sqrt(1.25) H, or when its crossing point falls
        off the emitter (a thin strip along the edges that face xp)."""
        a, b = float(np.linalg.norm(self.e1)), float(np.linalg.norm(self.e2))
        u1, u2 = self.e1 / a, self.e2 / b
        rel = xp - self.corner
        xs, xt = float(rel @ u1) / a, float(rel @ u2) / b
        height = float(self.corner[2] - xp[2])
        k = 0.002 / height
        rho2 = 1.25 * height * height
        # the inner integral's kinks: the disk's extent, where its boundary
        # crosses the emitter's edges and the strip edges, the strip switches
        cuts = [0.0, 1.0, xs - math.sqrt(rho2) / a, xs + math.sqrt(rho2) / a, -k * xs / (1 - k), (1 - k * xs) / (1 - k)]
        for tb in (0.0, 1.0, -k * xt / (1 - k), (1 - k * xt) / (1 - k)):
            rest = rho2 - ((tb - xt) * b) ** 2
            if rest >= 0:
                cuts += [xs - math.sqrt(rest) / a, xs + math.sqrt(rest) / a]
        cuts = sorted(c for c in cuts if 0.0 <= c <= 1.0)
        xg, wg = np.polynomial.legendre.leggauss(n)
        pts, wts = [], []
        for s0, s1 in zip(cuts[:-1], cuts[1:]):
            if s1 - s0 < 1e-15:
                continue
            for sn, sw in zip(0.5 * (s1 - s0) * (xg + 1) + s0, 0.5 * (s1 - s0) * wg):
                ivs = []
                rhs = rho2 - ((sn - xs) * a) ** 2
                if rhs >= 0:
                    h = math.sqrt(rhs) / b
                    ivs.append((max(0.0, xt - h), min(1.0, xt + h)))
                if 0.0 <= (1 - k) * sn + k * xs <= 1.0:
                    ivs.append((0.0, min(1.0, max(0.0, -k * xt / (1 - k)))))
                    ivs.append((min(1.0, max(0.0, (1 - k * xt) / (1 - k))), 1.0))
                else:
                    ivs.append((0.0, 1.0))
                for t0, t1 in _union(ivs):
                    tn = 0.5 * (t1 - t0) * (xg + 1) + t0
                    pts.append(self.corner + sn * self.e1 + tn[:, None] * self.e2)
                    wts.append(sw * 0.5 * (t1 - t0) * wg * a * b)
        return np.concatenate(pts), np.concatenate(wts)

    def _geometry(self, x, y):
        """Unit directions x -> y, distances, cosines at the plane and at the emitter."""
        v = y[None, :, :] - x[:, None, :]
        r = np.linalg.norm(v, axis=-1)
        w = v / r[..., None]
        return w, r, w[..., 2], w[..., 2]  # plane normal +z, emitter normal -z: both cosines = w_z

    def _nee_term(self, xp, wo, nomis, order):
        y, wa = self.visible_points(xp, order)
        area = float(np.linalg.norm(np.cross(self.e1, self.e2)))
        w, r, cs, cl = self._geometry(xp[None], y)
        w, r, cs, cl = w[0], r[0], cs[0], cl[0]
        f, bpdf = opaque_bsdf(w, wo[None, :], np.array([0.0, 0.0, 1.0]), self.rough, self.metallic, self.albedo)
        integrand = f * self.emission * (cs * cl / (r * r))[:, None]
        if not nomis:
            pdf_l = r * r / (cl * area)
            integrand = integrand * (pdf_l / (pdf_l + bpdf))[:, None]
        return integrand.T @ wa

    def expectation_depth1(self, cam, width, height, nomis=True, order=32):
        """E[L] of a depth-1 estimate: emission (none) + NEE at the plane over
        the emitter part its shadow rays reach, MIS-weighted against BSDF
        sampling when nomis is False."""
        x, wo = self.hits(cam, width, height)
        return np.stack([self._nee_term(x[p], wo[p], nomis, order) for p in range(len(x))])

    def expectation_depth2_mis(self, cam, width, height, order=24):
        """E[L] at depth 2 with MIS: the MIS-weighted NEE at the plane plus the
        emitter hit by the BSDF-sampled ray (which leaves from 0.001 above the
        plane), weighted by pdf_bsdf / (pdf_bsdf + M p_hat(z) / sumW), where
        sumW = p_hat(y)/pdf(y) comes from the NEE sample y of the same vertex,
        occluded or not (:595-625, :352-363)."""
        first = self.expectation_depth1(cam, width, height, nomis=False, order=order)
        x, wo = self.hits(cam, width, height)
        y, wa, area = self.light_points(order)
        n = np.array([0.0, 0.0, 1.0])
        out = np.zeros_like(first)
        le = self.emission
        for p in range(len(x)):
            # NEE samples y, seen from the vertex itself
            w, r, cs, cl = (v[0] for v in self._geometry(x[p:p + 1], y))
            fy, _ = opaque_bsdf(w, wo[p][None, :], n, self.rough, self.metallic, self.albedo)
            ph_y = (fy * le) @ LUM * cs
            pdf_y = r * r / (cl * area)
            # emitter points z hit by the continuation ray from x + 0.001 n
            wz, rz, csz, clz = (v[0] for v in self._geometry(x[p:p + 1] + 0.001 * n, y))
            fz, bz = opaque_bsdf(wz, wo[p][None, :], n, self.rough, self.metallic, self.albedo)
            ph_z = (fz * le) @ LUM * csz
            pl = ph_z[:, None] * pdf_y[None, :] / ph_y[None, :]
            weight = bz[:, None] / (bz[:, None] + pl)
            mean_w = weight @ (wa / area)
            integrand = fz * le * (csz * clz / (rz * rz) * mean_w)[:, None]
            out[p] = integrand.T @ wa
        return first + out


def _union(intervals):
    """Union of closed intervals as sorted, disjoint, non-empty intervals."""
    ivs = sorted((a, b) for a, b in intervals if b > a)
    out = []
    for a, b in ivs:
        if out and a <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], b))
        else:
            out.append((a, b))
    return out


# ---------------------------------------------------------------- furnace sphere
class FurnaceSphere:
    """A closed unit sphere (icosphere, smooth normals) whose walls all emit
    `emission` and reflect with one opaque BSDF; the camera is inside."""
    albedo = (0.7, 0.5, 0.3)
    metallic = 0.2
    rough = 0.7
    emission = np.array([0.5, 0.4, 0.3])

    def __init__(self, subdivisions=6):
        tris = scenes.icosphere(subdivisions)[:, [0, 2, 1]]  # inward winding: emitters face the interior
        normals = -tris / np.linalg.norm(tris, axis=2, keepdims=True)
        s = rm.Scene()
        s.add_triangles(tris, material_rows(len(tris), albedo=self.albedo, emission=tuple(self.emission),
                                            metallic=self.metallic, rough=self.rough), normals=normals)
        s.build_bvh()
        self.exported = s.export()

    @staticmethod
    def camera(width):
        return look_at((0.3, -0.2, 0.1), (-0.4, 0.6, -0.3), (width / 2) / math.tan(math.radians(30.0)))

    def view_cosines(self, cam, width, height):
        o, d = corner_rays(cam, width, height)
        b = np.sum(o * d, 1)
        t = -b + np.sqrt(b * b - (np.sum(o * o, 1) - 1.0))
        x = o + t[:, None] * d
        return np.sum(d * x, 1)  # cos between -d and the inward normal -x

    def series(self, mu_eval, depth, n_mu=64, n_phi=128):
        """sum_{k=0}^{depth} (T^k 1)(mu) per channel at each mu_eval, with
        (T g)(mu_o) = integral over the hemisphere of f(w_i, w_o) g(mu_i) mu_i dw_i."""
        xg, wg = np.polynomial.legendre.leggauss(n_mu)
        mu = 0.5 * (xg + 1.0)
        wmu = 0.5 * wg
        phi = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
        n = np.array([0.0, 0.0, 1.0])

        def kernel(mu_o):
            """K[o, j, c] = mu_j w_j integral over phi of f(mu_j, phi; mu_o)."""
            so = np.sqrt(1.0 - mu_o ** 2)
            wo = np.stack([so, np.zeros_like(so), mu_o], -1)[:, None, None, :]
            si = np.sqrt(1.0 - mu ** 2)
            wi = np.stack([si[:, None] * np.cos(phi)[None, :], si[:, None] * np.sin(phi)[None, :],
                           np.broadcast_to(mu[:, None], (n_mu, n_phi))], -1)[None]
            f, _ = opaque_bsdf(wi, wo, n, self.rough, self.metallic, self.albedo)
            return f.sum(axis=2) * (2.0 * np.pi / n_phi) * (mu * wmu)[None, :, None]

        k_nodes = kernel(mu)                     # (n_mu, n_mu, 3)
        k_eval = kernel(np.asarray(mu_eval, np.float64))
        g = np.ones((n_mu, 3))
        total = np.ones((len(mu_eval), 3))
        for _ in range(depth):
            total += np.einsum("ojc,jc->oc", k_eval, g)
            g = np.einsum("ojc,jc->oc", k_nodes, g)
        return total

    def expectation(self, cam, width, height, depth):
        mu = self.view_cosines(cam, width, height)
        return self.emission * self.series(mu, depth)


# ---------------------------------------------------------------- glass slab
def fresnel_unpolarised(cos_i, eta_i, eta_t):
    """Exact Fresnel reflectance from Snell's law and the Fresnel equations."""
    sin_t = eta_i / eta_t * np.sqrt(np.maximum(0.0, 1.0 - cos_i ** 2))
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t ** 2))
    rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t)
    rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t)
    return np.where(sin_t >= 1.0, 1.0, 0.5 * (rs * rs + rp * rp))


class GlassSlab:
    """A smooth glass slab (z in [1, 1.5], index 1.5) between an emitting
    floor (z = 0, facing up) and an emitting ceiling (z = 3, facing down); the
    camera sits between the slab and the ceiling, looking down at the slab."""
    ior = 1.5
    e_floor = np.array([2.0, 1.0, 0.5])
    e_ceiling = np.array([0.3, 0.6, 0.9])

    def __init__(self):
        s = rm.Scene()
        big, wall = 200.0, 100.0
        up = np.array([[0.0, 0.0, 1.0]] * 3)
        down = -up
        floor = scenes._quad((-wall, -wall, 0), (wall, -wall, 0), (wall, wall, 0), (-wall, wall, 0))
        ceiling = scenes._quad((-wall, -wall, 3), (-wall, wall, 3), (wall, wall, 3), (wall, -wall, 3))
        top = scenes._quad((-big, -big, 1.5), (big, -big, 1.5), (big, big, 1.5), (-big, big, 1.5))
        bottom = scenes._quad((-big, -big, 1.0), (-big, big, 1.0), (big, big, 1.0), (big, -big, 1.0))
        wall_mat = dict(albedo=(0.5, 0.5, 0.5), rough=0.5)
        s.add_triangles(floor, material_rows(2, emission=tuple(self.e_floor), **wall_mat), normals=np.array([up, up]))
        s.add_triangles(ceiling, material_rows(2, emission=tuple(self.e_ceiling), **wall_mat),
                        normals=np.array([down, down]))
        glass = material_rows(2, albedo=(1, 1, 1), ior=self.ior, rough=0.0)
        s.add_triangles(top, glass, normals=np.array([up, up]))
        s.add_triangles(bottom, glass, normals=np.array([down, down]))
        s.build_bvh()
        self.exported = s.export()

    @staticmethod
    def camera(width):
        return look_at((0.0, 0.0, 2.2), (0.0, 1.4, 1.0), (width / 2) / math.tan(math.radians(20.0)))

    def expectation(self, cam, width, height, depth):
        """Depth 2: ceiling seen by reflection, E_c F. Depth 3: plus the floor
        through both faces, E_f (1 - F)^2. (Deeper paths reach a wall only
        after one of its own bounces, whose BSDF enters.)"""
        _, d = corner_rays(cam, width, height)
        cos_t = -d[:, 2]
        assert np.all(cos_t > 0.3)
        f = fresnel_unpolarised(cos_t, 1.0, self.ior)[:, None]
        if depth == 2:
            return self.e_ceiling * f
        if depth == 3:
            return self.e_ceiling * f + self.e_floor * (1.0 - f) ** 2
        raise ValueError("closed form for depth 2 and 3 only")


# ---------------------------------------------------------------- hue shader
def hue_expected(pos):
    """Albedo the hue shader gives a hit at pos: h = mod(azimuth, 360) / 360
    with azimuth = atan(x/z) in degrees (+180 where z <= 0), passed to an
    hsv2rgb that takes degrees (so h in [0, 1) stays in its first sextant)
    and returns 0..255 (pathtrace.metal:26-63): 255 * hsv(h / 360, 0.8, 1)."""
    pos = np.asarray(pos, np.float64)
    az = np.degrees(np.arctan(pos[:, 0] / pos[:, 2])) + np.where(pos[:, 2] > 0, 0.0, 180.0)
    h = np.mod(az, 360.0) / 360.0
    return 255.0 * np.array([colorsys.hsv_to_rgb(v / 360.0, 0.8, 1.0) for v in h])


def hue_scene(azimuth_deg):
    """A lit quad of hue-shaded triangles centred at the given azimuth, 2 units
    from the origin, facing it; the camera at the origin looks at its centre."""
    a = math.radians(azimuth_deg)
    c = np.array([2.0 * math.sin(a), 0.0, 2.0 * math.cos(a)])
    nrm = -c / 2.0
    side = np.cross(nrm, [0.0, 1.0, 0.0])
    up = np.array([0.0, 1.0, 0.0])
    q = scenes._quad(c - side - up, c + side - up, c + side + up, c - side + up)
    if np.dot(np.cross(q[0, 1] - q[0, 0], q[0, 2] - q[0, 0]), nrm) < 0:
        q = q[:, [0, 2, 1]]
    s = rm.Scene()
    nn = np.broadcast_to(nrm, (2, 3, 3))
    s.add_triangles(q, material_rows(2, albedo=(0.5, 0.5, 0.5), shader=1), normals=nn)
    light = scenes._quad((-0.3, 3, -0.3), (0.3, 3, -0.3), (0.3, 3, 0.3), (-0.3, 3, 0.3))
    if np.cross(light[0, 1] - light[0, 0], light[0, 2] - light[0, 0])[1] > 0:
        light = light[:, [0, 2, 1]]
    s.add_triangles(light, material_rows(2, albedo=(0, 0, 0), emission=(5, 5, 5)))
    s.build_bvh()
    cam = look_at((0.0, 0.0, 0.0), c, 40.0, up=(0.0, 1.0, 0.0))
    return s.export(), cam


# ---------------------------------------------------------------- Monte Carlo checks
def check_mc(batches, expected, rel_sys=0.0, z_image=5.0, z_pixel=6.5, what=""):
    """batches (B, P, 3): per-batch pixel means from independent seeds;
    expected (P, 3). The image mean must agree within z_image standard errors
    (batch spread), every pixel within z_pixel of its own standard error (the
    batch spread, t-distributed with B - 1 degrees of freedom), and the RMS of
    the pixel z-scores stay near one; each error is widened by a relative
    systematic allowance rel_sys."""
    b = np.asarray(batches, np.float64)
    nb = len(b)
    mean = b.mean(0)
    img = b.mean(1)
    img_se = img.std(0, ddof=1) / math.sqrt(nb)
    exp_img = expected.mean(0)
    z = (img.mean(0) - exp_img) / np.sqrt(img_se ** 2 + (rel_sys * exp_img) ** 2)
    assert np.all(np.abs(z) < z_image), f"{what}: image mean {img.mean(0)} vs {exp_img} (z {z})"
    live = expected > 1e-6 * expected.max()
    se = b.std(0, ddof=1) / math.sqrt(nb)
    zp = np.abs(mean - expected)[live] / np.sqrt(se[live] ** 2 + (rel_sys * expected[live]) ** 2 + 1e-300)
    worst = float(zp.max())
    rms = float(np.sqrt(np.mean(zp ** 2)))
    assert worst < z_pixel and rms < 1.6, f"{what}: worst pixel z {worst:.2f}, RMS z {rms:.2f}"
    return {"z_image": z, "worst_pixel_z": worst, "rms_pixel_z": rms}


def seeds(n, base=1000):
    return [base + 7919 * i for i in range(n)]


# ---------------------------------------------------------------- denoiser inputs
def gbuffer(width, height, pos=(0.0, 0.0, 5.0), normal=(0.0, 0.0, 1.0), albedo=(0.5, 0.5, 0.5), kind=1):
    from raymotion_amd._capi import GBUFFER_DTYPE
    g = np.zeros(width * height, GBUFFER_DTYPE)
    g["pos"] = pos
    g["normal"] = normal
    g["albedo"] = albedo
    g["wo"] = (0.0, 0.0, -1.0)
    g["rough"] = 0.5
    g["ior"] = 1.0
    g["kind"] = kind
    return g


def constant_luma_colours(n, rng, luma=1.0):
    """n RGB colours of equal luminance and varied chroma."""
    u = np.array([1.0, 0.0, -0.2126 / 0.0722])
    v = np.cross(LUM, u)
    a = rng.uniform(-0.3, 0.3, n)
    b = rng.uniform(-0.05, 0.05, n)
    c = luma * np.ones((n, 3)) + a[:, None] * u / np.linalg.norm(u) * 0.2 + b[:, None] * v / np.linalg.norm(v)
    assert np.all(c > 0)
    return c


def box_filter(img, valid, step):
    """One a-trous pass with all edge-stopping weights equal to one: the mean
    over the 5x5 taps at stride 2^step that lie in the image and are not sky;
    sky pixels give zero."""
    h, w, _ = img.shape
    out = np.zeros_like(img)
    st = 1 << step
    for y in range(h):
        for x in range(w):
            if not valid[y, x]:
                continue
            acc = np.zeros(3)
            cnt = 0
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    yy, xx = y + dy * st, x + dx * st
                    if 0 <= yy < h and 0 <= xx < w and valid[yy, xx]:
                        acc += img[yy, xx]
                        cnt += 1
            out[y, x] = acc / cnt
    return out


# ---------------------------------------------------------------- ReSTIR merge
SPATIAL_SALT = 0x1B54A35A6F31B458  # low 64 bits of pathtrace.metal:937's 68-bit literal


def spatial_neighbours(seed, width, height, count, radius, accept):
    """For each pixel, the neighbours restir_spatial merges: count offsets
    (2 draws each) rounded as int(x + r * radius + 0.5) (truncation), skipped
    outside the image or on the pixel itself, and for `accept(id, nid)` one
    more draw (the selection) per merge. Returns a list of neighbour ids per pixel."""
    out = []
    for pid in range(width * height):
        x, y = pid % width, pid // width
        with np.errstate(over="ignore"):
            s = np.uint64(seed) ^ (np.uint64(pid + 1) * np_rng.GOLDEN) ^ np.uint64(SPATIAL_SALT)
        state = np_rng.splitmix_state(np.array([s], np.uint64))
        merged = []
        for _ in range(count):
            u, state = np_rng.draws(state, 2)
            rx = np.float32(u[0, 0]) * np.float32(2.0) - np.float32(1.0)
            ry = np.float32(u[0, 1]) * np.float32(2.0) - np.float32(1.0)
            fx = np.float32(np.float32(x) + rx * np.float32(radius)) + np.float32(0.5)
            fy = np.float32(np.float32(y) + ry * np.float32(radius)) + np.float32(0.5)
            nx, ny = int(np.trunc(fx)), int(np.trunc(fy))
            if nx < 0 or ny < 0 or nx >= width or ny >= height or (nx == x and ny == y):
                continue
            nid = ny * width + nx
            if not accept(pid, nid):
                continue
            merged.append(nid)
            _, state = np_rng.draws(state, 1)
        out.append(merged)
    return out
