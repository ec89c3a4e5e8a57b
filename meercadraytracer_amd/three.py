"""A small three.js-shaped scene graph, enough to describe the scenes the trace path renders.

The reference compiles real three.js objects (src/compile.ts). There is no three.js in Python, so
this module provides objects with the same names, defaults and matrix conventions (column-major
Matrix4, Euler XYZ rotation, sRGB hex colours decoded to linear as three.js's ColorManagement
does), and meercadraytracer_amd.compile turns them into the plain-data format. It covers meshes
(BufferGeometry with plane, box, icosphere and heightfield builders), the six mesh materials,
the five light types, perspective and orthographic cameras and fog.
"""
import math

import numpy as np

from . import abi

DEG2RAD = math.pi / 180


def srgb_to_linear(c):
    return c * 0.0773993808 if c < 0.04045 else (c * 0.9478672986 + 0.0521327014) ** 2.4


class Color:
    """THREE.Color: Color(r, g, b) is linear; Color(0xRRGGBB) is sRGB, decoded to linear."""

    def __init__(self, r=1.0, g=None, b=None):
        self.set(r, g, b)

    def set(self, r, g=None, b=None):
        if isinstance(r, Color):
            self.r, self.g, self.b = r.r, r.g, r.b
        elif g is None:
            h = int(r)
            self.r = srgb_to_linear(((h >> 16) & 255) / 255)
            self.g = srgb_to_linear(((h >> 8) & 255) / 255)
            self.b = srgb_to_linear((h & 255) / 255)
        else:
            self.r, self.g, self.b = float(r), float(g), float(b)
        return self


def _color(c, default=0xffffff):
    if c is None:
        return Color(default)
    return c if isinstance(c, Color) else Color(c)


# ---- Matrix4 helpers (column-major lists of 16) ---------------------------------------------
def compose(position, quaternion, scale):
    x, y, z, w = quaternion
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    sx, sy, sz = scale
    return [(1 - (yy + zz)) * sx, (xy + wz) * sx, (xz - wy) * sx, 0.0,
            (xy - wz) * sy, (1 - (xx + zz)) * sy, (yz + wx) * sy, 0.0,
            (xz + wy) * sz, (yz - wx) * sz, (1 - (xx + yy)) * sz, 0.0,
            position[0], position[1], position[2], 1.0]


def multiply4(a, b):
    out = [0.0] * 16
    for col in range(4):
        for row in range(4):
            out[col * 4 + row] = (a[row] * b[col * 4] + a[4 + row] * b[col * 4 + 1] +
                                  a[8 + row] * b[col * 4 + 2] + a[12 + row] * b[col * 4 + 3])
    return out


def invert4(m):
    a = np.array(m, dtype=np.float64).reshape(4, 4).T
    return list(np.linalg.inv(a).T.reshape(16))


def quat_from_euler(x, y, z):
    c1, c2, c3 = math.cos(x / 2), math.cos(y / 2), math.cos(z / 2)
    s1, s2, s3 = math.sin(x / 2), math.sin(y / 2), math.sin(z / 2)
    return (s1 * c2 * c3 + c1 * s2 * s3, c1 * s2 * c3 - s1 * c2 * s3,
            c1 * c2 * s3 + s1 * s2 * c3, c1 * c2 * c3 - s1 * s2 * s3)


def quat_from_rotation(te):
    m11, m12, m13 = te[0], te[4], te[8]
    m21, m22, m23 = te[1], te[5], te[9]
    m31, m32, m33 = te[2], te[6], te[10]
    trace = m11 + m22 + m33
    if trace > 0:
        s = 0.5 / math.sqrt(trace + 1.0)
        return ((m32 - m23) * s, (m13 - m31) * s, (m21 - m12) * s, 0.25 / s)
    if m11 > m22 and m11 > m33:
        s = 2.0 * math.sqrt(1.0 + m11 - m22 - m33)
        return (0.25 * s, (m12 + m21) / s, (m13 + m31) / s, (m32 - m23) / s)
    if m22 > m33:
        s = 2.0 * math.sqrt(1.0 + m22 - m11 - m33)
        return ((m12 + m21) / s, 0.25 * s, (m23 + m32) / s, (m13 - m31) / s)
    s = 2.0 * math.sqrt(1.0 + m33 - m11 - m22)
    return ((m13 + m31) / s, (m23 + m32) / s, 0.25 * s, (m21 - m12) / s)


def _norm(v):
    length = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / length, v[1] / length, v[2] / length] if length > 0 else list(v)


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


class Object3D:
    is_camera = False

    def __init__(self):
        self.position = [0.0, 0.0, 0.0]
        self.rotation = [0.0, 0.0, 0.0]
        self.scale = [1.0, 1.0, 1.0]
        self.quaternion = None  # set by lookAt; otherwise from rotation
        self.children = []
        self.parent = None
        self.visible = True
        self.castShadow = False
        self.receiveShadow = False
        self.matrixWorld = compose(self.position, (0, 0, 0, 1), self.scale)

    def add(self, *objects):
        for o in objects:
            o.parent = self
            self.children.append(o)
        return self

    def local_matrix(self):
        q = self.quaternion if self.quaternion is not None else quat_from_euler(*self.rotation)
        return compose(self.position, q, self.scale)

    def updateMatrixWorld(self):
        local = self.local_matrix()
        self.matrixWorld = local if self.parent is None else multiply4(self.parent.matrixWorld, local)
        for c in self.children:
            c.updateMatrixWorld()

    def lookAt(self, x, y, z):
        """Object3D.lookAt: cameras point -z at the target, other objects point +z."""
        self.updateMatrixWorld()
        pos = self.matrixWorld[12:15]
        eye, target = (pos, [x, y, z]) if self.is_camera else ([x, y, z], pos)
        zc = [eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]]
        if zc == [0, 0, 0]:
            zc = [0, 0, 1]
        zc = _norm(zc)
        up = [0.0, 1.0, 0.0]
        xc = _cross(up, zc)
        if xc == [0, 0, 0] or sum(v * v for v in xc) == 0:
            zc[2] += 0.0001 if abs(up[2]) == 1 else 0
            zc[0] += 0.0001 if abs(up[2]) != 1 else 0
            zc = _norm(zc)
            xc = _cross(up, zc)
        xc = _norm(xc)
        yc = _cross(zc, xc)
        te = [xc[0], xc[1], xc[2], 0, yc[0], yc[1], yc[2], 0, zc[0], zc[1], zc[2], 0, 0, 0, 0, 1]
        self.quaternion = quat_from_rotation(te)
        self.updateMatrixWorld()


class Scene(Object3D):
    def __init__(self):
        super().__init__()
        self.background = None  # None, a Color or a texture
        self.backgroundIntensity = 1.0
        self.backgroundRotation = None  # Euler (x, y, z) or None
        self.fog = None
        self.environment = None
        self.environmentIntensity = 1.0
        self.environmentRotation = None


# ---- textures --------------------------------------------------------------------------------
UVMapping, CubeReflectionMapping, CubeRefractionMapping = 300, 301, 302
EquirectangularReflectionMapping, EquirectangularRefractionMapping = 303, 304
UnsignedByteType, FloatType, HalfFloatType = 1009, 1015, 1016
NoColorSpace, SRGBColorSpace = '', 'srgb'


class DataTexture:
    """THREE.DataTexture with RGBA data, rows bottom to top as three.js uploads them (flipY false).
    type: UnsignedByteType (uint8), FloatType (float32) or HalfFloatType (uint16 half floats)."""
    isTexture = True
    isCubeTexture = False

    def __init__(self, data, width, height, type=UnsignedByteType):
        self.image = {'data': np.asarray(data), 'width': int(width), 'height': int(height)}
        self.type = type
        self.colorSpace = NoColorSpace
        self.flipY = False
        self.wrapS = self.wrapT = abi.ClampToEdgeWrapping
        self.mapping = UVMapping
        self.offset = [0.0, 0.0]
        self.repeat = [1.0, 1.0]
        self.center = [0.0, 0.0]
        self.rotation = 0.0
        self.matrixAutoUpdate = True
        self.matrix = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
        self.version = 0

    def updateMatrix(self):  # Texture.updateMatrix -> Matrix3.setUvTransform, column-major
        tx, ty = self.offset
        sx, sy = self.repeat
        cx, cy = self.center
        c, s = math.cos(self.rotation), math.sin(self.rotation)
        self.matrix = [sx * c, -sy * s, 0.0, sx * s, sy * c, 0.0,
                       -sx * (c * cx + s * cy) + cx + tx, -sy * (-s * cx + c * cy) + cy + ty, 1.0]


class CubeTexture:
    """THREE.CubeTexture from six RGBA DataTexture-like images (+x, -x, +y, -y, +z, -z)."""
    isTexture = True
    isCubeTexture = True
    isRenderTargetTexture = False

    def __init__(self, faces, type=UnsignedByteType):
        self.image = list(faces)
        self.type = type
        self.colorSpace = NoColorSpace
        self.mapping = CubeReflectionMapping
        self.wrapS = self.wrapT = abi.ClampToEdgeWrapping
        self.version = 0


# ---- geometry --------------------------------------------------------------------------------
class BufferGeometry:
    def __init__(self, position, normal=None, uv=None, index=None, color=None):
        self.position = np.asarray(position, dtype=np.float32).reshape(-1, 3)
        self.normal = None if normal is None else np.asarray(normal, dtype=np.float32).reshape(-1, 3)
        self.uv = None if uv is None else np.asarray(uv, dtype=np.float32).reshape(-1, 2)
        self.index = None if index is None else np.asarray(index, dtype=np.int64).reshape(-1)
        self.color = None if color is None else np.asarray(color, dtype=np.float32)
        self.color_size = 0 if color is None else self.color.shape[1]

    def setAttribute(self, name, values, item_size):
        arr = np.asarray(values, dtype=np.float32).reshape(-1, item_size)
        if name == 'color':
            self.color, self.color_size = arr, item_size
        else:
            setattr(self, name, arr)


def PlaneGeometry(width=1.0, height=1.0, wseg=1, hseg=1):
    """THREE.PlaneGeometry: a grid in z = 0 facing +z, rows from +y down."""
    gx, gy = int(wseg), int(hseg)
    seg_w, seg_h = width / gx, height / gy
    pos, nrm, uv, idx = [], [], [], []
    for iy in range(gy + 1):
        y = iy * seg_h - height / 2
        for ix in range(gx + 1):
            x = ix * seg_w - width / 2
            pos.append([x, -y, 0.0])
            nrm.append([0.0, 0.0, 1.0])
            uv.append([ix / gx, 1 - iy / gy])
    for iy in range(gy):
        for ix in range(gx):
            a = ix + (gx + 1) * iy
            b = ix + (gx + 1) * (iy + 1)
            c = (ix + 1) + (gx + 1) * (iy + 1)
            d = (ix + 1) + (gx + 1) * iy
            idx += [a, b, d, b, c, d]
    return BufferGeometry(pos, nrm, uv, idx)


def BoxGeometry(width=1.0, height=1.0, depth=1.0):
    """A box centred on the origin: 6 faces of 2 triangles, counter-clockwise seen from outside."""
    hw, hh, hd = width / 2, height / 2, depth / 2
    faces = [  # (normal, u axis, v axis) with u x v = normal
        ((1, 0, 0), (0, 0, -1), (0, 1, 0)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
        ((0, 1, 0), (1, 0, 0), (0, 0, -1)), ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
        ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (-1, 0, 0), (0, 1, 0)),
    ]
    half = np.array([hw, hh, hd])
    pos, nrm, uv, idx = [], [], [], []
    for n, u, v in faces:
        n, u, v = np.array(n, float), np.array(u, float), np.array(v, float)
        base = len(pos)
        for su, sv in ((-1, 1), (1, 1), (-1, -1), (1, -1)):
            p = (n + su * u + sv * v) * half
            pos.append(p)
            nrm.append(n)
            uv.append([(su + 1) / 2, (sv + 1) / 2])
        idx += [base, base + 2, base + 1, base + 2, base + 3, base + 1]
    return BufferGeometry(pos, nrm, uv, idx)


def IcosphereGeometry(radius=1.0, subdivisions=3):
    """Icosahedron refined by recursive 4-way subdivision: 20 * 4^s triangles, vertices on the
    sphere, smooth normals (the normalised position)."""
    t = (1 + math.sqrt(5)) / 2
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [list(np.array(v) / np.linalg.norm(v)) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = np.array(verts, dtype=np.float64)
    f = np.array(faces, dtype=np.int64)
    for _ in range(int(subdivisions)):
        edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)  # mid of (0,1), (1,2), (2,0)
        v = np.concatenate([v, mids])
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    pos = v * radius
    uv = np.zeros((len(v), 2))
    return BufferGeometry(pos, v, uv, f.reshape(-1))


def HeightfieldGeometry(heights, size_x, size_z):
    """A terrain grid in the xz plane, y = heights[row, col], facing +y. (rows-1)*(cols-1)*2
    triangles with smooth normals from central differences."""
    rows, cols = heights.shape
    xs = np.linspace(-size_x / 2, size_x / 2, cols)
    zs = np.linspace(-size_z / 2, size_z / 2, rows)
    X, Z = np.meshgrid(xs, zs)
    pos = np.stack([X, heights, Z], -1).reshape(-1, 3)
    dx = xs[1] - xs[0]
    dz = zs[1] - zs[0]
    gx = np.gradient(heights, dx, axis=1)
    gz = np.gradient(heights, dz, axis=0)
    nrm = np.stack([-gx, np.ones_like(heights), -gz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack([(X - xs[0]) / size_x, (Z - zs[0]) / size_z], -1).reshape(-1, 2)
    r, c = np.meshgrid(np.arange(rows - 1), np.arange(cols - 1), indexing='ij')
    a = (r * cols + c).reshape(-1)
    b = a + cols
    d = a + 1
    e = b + 1
    idx = np.stack([a, b, d, b, e, d], 1).reshape(-1)  # counter-clockwise seen from +y
    return BufferGeometry(pos, nrm.reshape(-1, 3), uv, idx)


# ---- materials -----------------------------------------------------------------------------
class Material:
    isMeshStandardMaterial = isMeshPhysicalMaterial = isMeshPhongMaterial = False
    isMeshLambertMaterial = isMeshToonMaterial = isMeshBasicMaterial = False

    def __init__(self, **kw):
        self.side = abi.FrontSide
        self.opacity = 1.0
        self.transparent = False
        self.alphaTest = 0.0
        self.vertexColors = False
        self.flatShading = False
        self.toneMapped = True
        self.visible = True
        self.wireframe = False
        self.color = Color(0xffffff)
        self.emissive = None
        self.emissiveIntensity = 1.0
        self._defaults()
        for k, v in kw.items():
            if k in ('color', 'emissive', 'specular', 'specularColor', 'attenuationColor'):
                v = _color(v)
            setattr(self, k, v)

    def _defaults(self):
        pass


class MeshStandardMaterial(Material):
    isMeshStandardMaterial = True

    def _defaults(self):
        self.roughness = 1.0
        self.metalness = 0.0
        self.emissive = Color(0x000000)
        self.envMapIntensity = 1.0


class MeshPhysicalMaterial(MeshStandardMaterial):
    isMeshPhysicalMaterial = True

    def _defaults(self):
        super()._defaults()
        self.ior = 1.5
        self.specularIntensity = 1.0
        self.specularColor = Color(1, 1, 1)
        self.transmission = 0.0
        self.thickness = 0.0
        self.attenuationDistance = math.inf
        self.attenuationColor = Color(1, 1, 1)


class MeshPhongMaterial(Material):
    isMeshPhongMaterial = True

    def _defaults(self):
        self.specular = Color(0x111111)
        self.shininess = 30.0
        self.emissive = Color(0x000000)


class MeshLambertMaterial(Material):
    isMeshLambertMaterial = True

    def _defaults(self):
        self.emissive = Color(0x000000)


class MeshToonMaterial(Material):
    isMeshToonMaterial = True

    def _defaults(self):
        self.emissive = Color(0x000000)


class MeshBasicMaterial(Material):
    isMeshBasicMaterial = True


class Mesh(Object3D):
    def __init__(self, geometry, material):
        super().__init__()
        self.geometry = geometry
        self.material = material


# ---- lights ----------------------------------------------------------------------------------
class Light(Object3D):
    kind = None

    def __init__(self, color=0xffffff, intensity=1.0):
        super().__init__()
        self.color = _color(color)
        self.intensity = float(intensity)


class AmbientLight(Light):
    kind = 'ambient'


class HemisphereLight(Light):
    kind = 'hemisphere'

    def __init__(self, sky=0xffffff, ground=0xffffff, intensity=1.0):
        super().__init__(sky, intensity)
        self.groundColor = _color(ground)
        self.position = [0.0, 1.0, 0.0]


class DirectionalLight(Light):
    kind = 'directional'

    def __init__(self, color=0xffffff, intensity=1.0):
        super().__init__(color, intensity)
        self.position = [0.0, 1.0, 0.0]
        self.target = Object3D()


class PointLight(Light):
    kind = 'point'

    def __init__(self, color=0xffffff, intensity=1.0, distance=0.0, decay=2.0):
        super().__init__(color, intensity)
        self.distance = float(distance)
        self.decay = float(decay)


class SpotLight(Light):
    kind = 'spot'

    def __init__(self, color=0xffffff, intensity=1.0, distance=0.0, angle=math.pi / 3,
                 penumbra=0.0, decay=2.0):
        super().__init__(color, intensity)
        self.position = [0.0, 1.0, 0.0]
        self.distance = float(distance)
        self.angle = float(angle)
        self.penumbra = float(penumbra)
        self.decay = float(decay)
        self.target = Object3D()


# ---- cameras ---------------------------------------------------------------------------------
class Camera(Object3D):
    is_camera = True
    isOrthographicCamera = False

    def projection(self):
        raise NotImplementedError


class PerspectiveCamera(Camera):
    def __init__(self, fov=50.0, aspect=1.0, near=0.1, far=2000.0):
        super().__init__()
        self.fov, self.aspect, self.near, self.far, self.zoom = fov, aspect, near, far, 1.0

    def projection(self):
        top = self.near * math.tan(DEG2RAD * 0.5 * self.fov) / self.zoom
        height = 2 * top
        width = self.aspect * height
        left = -0.5 * width
        right, bottom = left + width, top - height
        n, f = self.near, self.far
        x = 2 * n / (right - left)
        y = 2 * n / (top - bottom)
        a = (right + left) / (right - left)
        b = (top + bottom) / (top - bottom)
        c = -(f + n) / (f - n)
        d = (-2 * f * n) / (f - n)
        return [x, 0, 0, 0, 0, y, 0, 0, a, b, c, -1, 0, 0, d, 0]


class OrthographicCamera(Camera):
    isOrthographicCamera = True

    def __init__(self, left=-1.0, right=1.0, top=1.0, bottom=-1.0, near=0.1, far=2000.0):
        super().__init__()
        self.left, self.right, self.top, self.bottom = left, right, top, bottom
        self.near, self.far, self.zoom = near, far, 1.0

    def projection(self):
        dx = (self.right - self.left) / (2 * self.zoom)
        dy = (self.top - self.bottom) / (2 * self.zoom)
        cx = (self.right + self.left) / 2
        cy = (self.top + self.bottom) / 2
        left, right, top, bottom = cx - dx, cx + dx, cy + dy, cy - dy
        w = 1.0 / (right - left)
        h = 1.0 / (top - bottom)
        p = 1.0 / (self.far - self.near)
        x = (right + left) * w
        y = (top + bottom) * h
        z = (self.far + self.near) * p
        return [2 * w, 0, 0, 0, 0, 2 * h, 0, 0, 0, 0, -2 * p, 0, -x, -y, -z, 1]


class Fog:
    def __init__(self, color, near=1.0, far=1000.0):
        self.isFog, self.isFogExp2 = True, False
        self.color, self.near, self.far = _color(color), near, far


class FogExp2:
    def __init__(self, color, density=0.00025):
        self.isFog, self.isFogExp2 = False, True
        self.color, self.density = _color(color), density
