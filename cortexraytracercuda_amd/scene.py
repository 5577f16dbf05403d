"""Host-side scene description mirroring the reference's sceneConfig semantics.

The reference links materials and textures by name (core/src/scenes/scene_config.h:452-460,
resolved on the device by strcmp in core/src/gpu.cu:132-474). `Scene` keeps the same vocabulary
(lambertian/metal/dielectric materials, solid/checker/image textures, directional area lights,
spheres, quads, OBJ-style meshes, instances) but resolves names to indices on the host and
produces the flat C-ABI descriptor of include/cortex_rt.h.
"""
import ctypes as C

import numpy as np

from . import _abi as A


def _f3(v):
    return A.f3(*[float(x) for x in v])


class Scene:
    def __init__(self):
        self.camera = dict(lookfrom=(0.0, 0.0, 1.0), lookat=(0.0, 0.0, 0.0), vup=(0.0, 1.0, 0.0), vfov=70.0,
                           aspect_ratio=1.0, aperture=0.0, focus_dist=10.0, time0=0.0, time1=1.0)
        self.background = (0.0, 0.0, 0.0)
        self.materials = []      # dicts
        self.material_names = {}
        self.textures = []
        self.texture_names = {}
        self.lights = []
        self.sphere_lights = []  # omni and spot lights
        self.spheres = []
        self.quads = []
        self.shapes = []         # disks, cylinders, cones
        self.meshes = []         # dicts of numpy arrays
        self.instances = []

    # ---------------------------------------------------------------- camera / image
    def set_camera(self, lookfrom, lookat, vup=(0, 1, 0), vfov=70.0, aspect_ratio=1.0, aperture=0.0, focus_dist=10.0,
                   time0=0.0, time1=1.0, orthographic=False, ortho_height=1.0):
        """Perspective camera, or with orthographic=True parallel rays from a viewport of height
        2 * ortho_height centred at lookfrom (vfov and focus_dist unused)."""
        self.camera = dict(lookfrom=tuple(lookfrom), lookat=tuple(lookat), vup=tuple(vup), vfov=float(vfov),
                           aspect_ratio=float(aspect_ratio), aperture=float(aperture), focus_dist=float(focus_dist),
                           time0=float(time0), time1=float(time1), orthographic=bool(orthographic),
                           ortho_height=float(ortho_height))
        return self

    # ---------------------------------------------------------------- textures
    def _add_tex(self, name, d):
        idx = len(self.textures)
        self.textures.append(d)
        if name:
            self.texture_names[name] = idx
        return idx

    def solid_texture(self, name, rgb):
        return self._add_tex(name, dict(type=A.CRT_TEX_SOLID, color=tuple(rgb)))

    def checker_texture(self, name, scale, even, odd):
        """even/odd: texture names/indices or rgb triples (checker_texture.cuh:23-31)."""
        def sub(x, tag):
            if isinstance(x, str):
                return self.texture_names[x]
            if isinstance(x, (int, np.integer)):
                return int(x)
            return self.solid_texture(None, x)
        e, o = sub(even, "even"), sub(odd, "odd")
        return self._add_tex(name, dict(type=A.CRT_TEX_CHECKER, scale=float(scale), even=e, odd=o))

    def image_texture(self, name, pixels):
        """pixels: uint8 array (H, W, 3|4), row 0 at the top."""
        px = np.ascontiguousarray(pixels, dtype=np.uint8)
        if px.ndim != 3 or px.shape[2] not in (3, 4):
            raise ValueError("image texture must be HxWx3 or HxWx4 uint8")
        return self._add_tex(name, dict(type=A.CRT_TEX_IMAGE, pixels=px))

    # ---------------------------------------------------------------- materials
    def _add_mat(self, name, d):
        idx = len(self.materials)
        self.materials.append(d)
        if name:
            self.material_names[name] = idx
        return idx

    def _tex_index(self, texture):
        if texture is None:
            return -1
        if isinstance(texture, str):
            return self.texture_names[texture]
        return int(texture)

    def lambertian(self, name, rgb=(0.5, 0.5, 0.5), texture=None):
        return self._add_mat(name, dict(type=A.CRT_MAT_LAMBERTIAN, albedo=tuple(rgb), texture=self._tex_index(texture)))

    def metal(self, name, rgb, fuzz=0.0, texture=None):
        return self._add_mat(name, dict(type=A.CRT_MAT_METAL, albedo=tuple(rgb), fuzz=float(fuzz),
                                        texture=self._tex_index(texture)))

    def isotropic(self, name, rgb=(0.5, 0.5, 0.5), texture=None):
        """isotropic.cuh: albedo from the texture (or rgb), scattered over the whole sphere."""
        return self._add_mat(name, dict(type=A.CRT_MAT_ISOTROPIC, albedo=tuple(rgb), texture=self._tex_index(texture)))

    def oren_nayar(self, name, rgb=(0.5, 0.5, 0.5), roughness=0.0, albedo_temp=1.0):
        """oren_nayar.cuh with a solid albedo (as load_scene creates it): attenuation from light 0's
        colour and intensity, albedo_temp / pi and the roughness terms."""
        return self._add_mat(name, dict(type=A.CRT_MAT_OREN_NAYAR, albedo=tuple(rgb), fuzz=float(roughness),
                                        ior=float(albedo_temp)))

    def dielectric(self, name, ior):
        return self._add_mat(name, dict(type=A.CRT_MAT_DIELECTRIC, albedo=(1.0, 1.0, 1.0), ior=float(ior)))

    def diffuse_light(self, name, rgb, intensity=1.0, one_sided=False):
        return self._add_mat(name, dict(type=A.CRT_MAT_DIFFUSE_LIGHT, albedo=tuple(rgb), intensity=float(intensity),
                                        flags=A.CRT_MATF_ONE_SIDED if one_sided else 0))

    def _mat_index(self, m):
        if isinstance(m, str):
            return self.material_names[m]
        return int(m)

    # ---------------------------------------------------------------- geometry
    def area_light(self, center, u, v, rgb=(1, 1, 1), intensity=1.0, invisible=True):
        self.lights.append(dict(center=tuple(center), u=tuple(u), v=tuple(v), rgb=tuple(rgb), intensity=float(intensity),
                                invisible=bool(invisible)))
        return len(self.lights) - 1

    def omni_light(self, center, radius, rgb=(1, 1, 1), intensity=1.0, invisible=True):
        """Spherical light emitting rgb * intensity from both faces (omni_light.cuh)."""
        self.sphere_lights.append(dict(kind=A.CRT_LIGHT_OMNI, center=tuple(center), radius=float(radius), rgb=tuple(rgb),
                                       intensity=float(intensity), invisible=bool(invisible), direction=(0.0, 0.0, -1.0),
                                       cutoff_deg=0.0, falloff=0.0))
        return len(self.sphere_lights) - 1

    def spot_light(self, center, direction, cutoff_deg, falloff, radius, rgb=(1, 1, 1), intensity=1.0, invisible=True):
        """Spherical spot light (spot_light.cuh, diffuse_spot_light.cuh)."""
        self.sphere_lights.append(dict(kind=A.CRT_LIGHT_SPOT, center=tuple(center), radius=float(radius), rgb=tuple(rgb),
                                       intensity=float(intensity), invisible=bool(invisible), direction=tuple(direction),
                                       cutoff_deg=float(cutoff_deg), falloff=float(falloff)))
        return len(self.sphere_lights) - 1

    def sphere(self, center, radius, material, mapping=(1, 1, 0, 0, 1, 1)):
        self.spheres.append(dict(center=tuple(center), radius=float(radius), material=self._mat_index(material),
                                 mapping=tuple(mapping)))
        return len(self.spheres) - 1

    def quad(self, corner, u, v, material):
        self.quads.append(dict(corner=tuple(corner), u=tuple(u), v=tuple(v), material=self._mat_index(material)))
        return len(self.quads) - 1

    def _shape(self, kind, center, radius, height, material, mapping, translate, rotate_deg, scale):
        has = ((A.CRT_XF_TRANSLATE if translate is not None else 0) | (A.CRT_XF_ROTATE if rotate_deg is not None else 0) |
               (A.CRT_XF_SCALE if scale is not None else 0))
        self.shapes.append(dict(kind=kind, center=tuple(center), radius=float(radius), height=float(height),
                                material=self._mat_index(material), mapping=tuple(mapping), has=has,
                                translate=tuple(translate or (0, 0, 0)), rotate_deg=tuple(rotate_deg or (0, 0, 0)),
                                scale=tuple(scale or (1, 1, 1))))
        return len(self.shapes) - 1

    def disk(self, center, radius, height, material, mapping=(1, 1, 0, 0, 1, 1), translate=None, rotate_deg=None,
             scale=None):
        """Disk of `radius` in the plane y = center.y, facing +y (disk.cuh); `height` sizes its box.
        translate / rotate_deg / scale place it as the reference's wrappers do."""
        return self._shape(A.CRT_SHAPE_DISK, center, radius, height, material, mapping, translate, rotate_deg, scale)

    def cylinder(self, center, radius, height, material, mapping=(1, 1, 0, 0, 1, 1), translate=None, rotate_deg=None,
                 scale=None):
        """Open vertical tube from the base centre up `height` (cylinder.cuh)."""
        return self._shape(A.CRT_SHAPE_CYLINDER, center, radius, height, material, mapping, translate, rotate_deg, scale)

    def cone(self, center, radius, height, material, mapping=(1, 1, 0, 0, 1, 1), translate=None, rotate_deg=None,
             scale=None):
        """Cone side, base circle at center.y, apex `height` above (cone.cuh)."""
        return self._shape(A.CRT_SHAPE_CONE, center, radius, height, material, mapping, translate, rotate_deg, scale)

    def mesh(self, positions, indices, material, normals=None, uvs=None, tri_materials=None):
        m = dict(positions=np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3),
                 indices=np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3),
                 material=self._mat_index(material),
                 normals=None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3),
                 uvs=None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2),
                 tri_materials=None if tri_materials is None else np.ascontiguousarray(tri_materials, dtype=np.uint32))
        self.meshes.append(m)
        return len(self.meshes) - 1

    def load_obj(self, path, material, use_smoothing=True, transform=None):
        """Adds an OBJ mesh read by the library's loader (crt_obj_load). use_smoothing=False keeps
        flat normals like the reference's objMesh use_smoothing flag; `transform` (3x4) places it."""
        lib = A.load()
        m = A.crt_obj_mesh()
        A.check(lib.crt_obj_load(str(path).encode(), C.byref(m)))
        try:
            nv, nt = m.num_vertices, m.num_triangles
            pos = np.ctypeslib.as_array(m.positions, (nv * 3,)).reshape(-1, 3).copy() if nv else np.zeros((0, 3), np.float32)
            idx = np.ctypeslib.as_array(m.indices, (nt * 3,)).reshape(-1, 3).copy() if nt else np.zeros((0, 3), np.uint32)
            nrm = np.ctypeslib.as_array(m.normals, (nv * 3,)).reshape(-1, 3).copy() if m.normals else None
            uv = np.ctypeslib.as_array(m.uvs, (nv * 2,)).reshape(-1, 2).copy() if m.uvs else None
        finally:
            lib.crt_obj_free(C.byref(m))
        if transform is not None:
            x = np.asarray(transform, np.float64).reshape(3, 4)
            pos = (pos.astype(np.float64) @ x[:, :3].T + x[:, 3]).astype(np.float32)
            if nrm is not None:
                nrm = (nrm.astype(np.float64) @ np.linalg.inv(x[:, :3])).astype(np.float32)
        return self.mesh(pos, idx, material, normals=nrm if use_smoothing else None, uvs=uv)

    def instance(self, mesh, object_to_world, material_override=-1):
        x = np.asarray(object_to_world, dtype=np.float32).reshape(-1)
        if x.size == 16:
            x = x[:12]
        if x.size != 12:
            raise ValueError("object_to_world must be 3x4 or 4x4")
        mo = -1 if material_override is None or material_override == -1 else self._mat_index(material_override)
        self.instances.append(dict(mesh=int(mesh), xf=x.copy(), material_override=mo))
        return len(self.instances) - 1

    @property
    def num_triangles(self):
        if self.instances:
            return sum(int(self.meshes[i["mesh"]]["indices"].shape[0]) for i in self.instances)
        return sum(int(m["indices"].shape[0]) for m in self.meshes)

    @property
    def num_world_prims(self):
        return (len(self.lights) + len(self.sphere_lights) + len(self.spheres) + len(self.quads) + len(self.shapes) +
                sum(int(m["indices"].shape[0]) for m in self.meshes))

    # ---------------------------------------------------------------- C descriptor
    def to_desc(self):
        """Returns (crt_scene_desc, keepalive). The keepalive list owns every buffer the
        descriptor points to and must outlive the call that consumes the descriptor."""
        keep = []

        def arr(ctype, items):
            n = len(items)
            if n == 0:
                return None, 0
            a = (ctype * n)(*items)
            keep.append(a)
            return C.cast(a, C.POINTER(ctype)), n

        d = A.crt_scene_desc()
        d.abi_version = A.CRT_ABI_VERSION
        c = self.camera
        d.camera = A.crt_camera(_f3(c["lookfrom"]), _f3(c["lookat"]), _f3(c["vup"]), c["vfov"], c["aspect_ratio"],
                                c["aperture"], c["focus_dist"], c["time0"], c["time1"], 1 if c.get("orthographic") else 0,
                                c.get("ortho_height", 1.0))
        d.background = _f3(self.background)
        mats = []
        for m in self.materials:
            mats.append(A.crt_material(m["type"], m.get("texture", -1), _f3(m.get("albedo", (0, 0, 0))), m.get("fuzz", 0.0),
                                       m.get("ior", 1.0), m.get("intensity", 1.0), m.get("flags", 0)))
        d.materials, d.num_materials = arr(A.crt_material, mats)
        texs = []
        for t in self.textures:
            ct = A.crt_texture()
            ct.type = t["type"]
            ct.even = -1
            ct.odd = -1
            if t["type"] == A.CRT_TEX_SOLID:
                ct.color = _f3(t["color"])
            elif t["type"] == A.CRT_TEX_CHECKER:
                ct.scale = t["scale"]
                ct.even = t["even"]
                ct.odd = t["odd"]
            else:
                px = t["pixels"]
                keep.append(px)
                ct.height, ct.width, ct.channels = px.shape
                ct.pixels = px.ctypes.data_as(C.POINTER(C.c_uint8))
            texs.append(ct)
        d.textures, d.num_textures = arr(A.crt_texture, texs)
        d.lights, d.num_lights = arr(A.crt_area_light, [
            A.crt_area_light(_f3(l["center"]), _f3(l["u"]), _f3(l["v"]), _f3(l["rgb"]), l["intensity"],
                             1 if l["invisible"] else 0) for l in self.lights])
        d.spheres, d.num_spheres = arr(A.crt_sphere, [
            A.crt_sphere(_f3(s["center"]), s["radius"], s["material"], A.crt_uvmap(*[float(x) for x in s["mapping"]]))
            for s in self.spheres])
        d.quads, d.num_quads = arr(A.crt_quad, [
            A.crt_quad(_f3(q["corner"]), _f3(q["u"]), _f3(q["v"]), q["material"]) for q in self.quads])
        d.shapes, d.num_shapes = arr(A.crt_shape, [
            A.crt_shape(h["kind"], _f3(h["center"]), h["radius"], h["height"], h["material"],
                        A.crt_uvmap(*[float(x) for x in h["mapping"]]),
                        A.crt_transform(h["has"], _f3(h["translate"]), _f3(h["rotate_deg"]), _f3(h["scale"])))
            for h in self.shapes])
        meshes = []
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_uint32)
        for m in self.meshes:
            cm = A.crt_mesh()
            keep.extend([m["positions"], m["indices"]])
            cm.positions = m["positions"].ctypes.data_as(fp)
            cm.num_vertices = m["positions"].shape[0]
            cm.indices = m["indices"].ctypes.data_as(up)
            cm.num_triangles = m["indices"].shape[0]
            if m["normals"] is not None:
                keep.append(m["normals"])
                cm.normals = m["normals"].ctypes.data_as(fp)
            if m["uvs"] is not None:
                keep.append(m["uvs"])
                cm.uvs = m["uvs"].ctypes.data_as(fp)
            cm.material = m["material"]
            if m["tri_materials"] is not None:
                keep.append(m["tri_materials"])
                cm.tri_materials = m["tri_materials"].ctypes.data_as(up)
            meshes.append(cm)
        d.meshes, d.num_meshes = arr(A.crt_mesh, meshes)
        insts = []
        for i in self.instances:
            ci = A.crt_instance()
            ci.mesh = i["mesh"]
            ci.object_to_world = (C.c_float * 12)(*[float(x) for x in i["xf"]])
            ci.material_override = i["material_override"]
            insts.append(ci)
        d.instances, d.num_instances = arr(A.crt_instance, insts)
        d.sphere_lights, d.num_sphere_lights = arr(A.crt_sphere_light, [
            A.crt_sphere_light(l["kind"], _f3(l["center"]), l["radius"], _f3(l["rgb"]), l["intensity"],
                               1 if l["invisible"] else 0, _f3(l["direction"]), l["cutoff_deg"], l["falloff"])
            for l in self.sphere_lights])
        keep.append(d)
        return d, keep
