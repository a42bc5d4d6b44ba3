"""ORACLE - test infrastructure only.

The reference's pbrt-v4 reader and scene conversion restated in Python, as
the checker of phx_pbrt_load:
  parse          parse_pbrt_file (pbrt_parser.cpp:233-637): tokens (:19-94),
                 parameters (:104-160), named spectra (:170-183), Include /
                 Import splicing (:201-231), the directives
  read_ply       load_ply (ply_reader.cpp:64-251)
  read_pfm       load_pfm (image_loader.cpp:19-77); the other image formats
                 are not restated (the tests use PFM files)
  convert        convert_pbrt_scene (pbrt_to_photon.cpp:58-461), returning
                 the arrays of the converted scene, its camera and settings
Float arithmetic is float32 in the reference's operation order; strtof,
sqrtf, sinf, cosf come from the C library through ctypes, as the C++ code
calls them.
"""
import ctypes as C
import ctypes.util
import os
import struct

import numpy as np

f32 = np.float32
_libc = C.CDLL(ctypes.util.find_library("c"))
_libm = C.CDLL(ctypes.util.find_library("m"))
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
for _fn in ("sqrtf", "sinf", "cosf"):
    getattr(_libm, _fn).restype = C.c_float
    getattr(_libm, _fn).argtypes = [C.c_float]


def _strtof(s):
    return f32(_libc.strtof(s.encode(), None))


def _sqrt(x):
    return f32(_libm.sqrtf(float(x)))


def _sin(x):
    return f32(_libm.sinf(float(x)))


def _cos(x):
    return f32(_libm.cosf(float(x)))


# ---- tokens (pbrt_parser.cpp:19-94) -----------------------------------------------
STR, NUM, OPEN, CLOSE, WORD, END = range(6)


class Lexer:
    def __init__(self, text):
        self.s = text
        self.i = 0

    def next(self):
        s, n = self.s, len(self.s)
        while True:
            while self.i < n:
                c = s[self.i]
                if c.isspace():
                    self.i += 1
                elif c == "#":
                    while self.i < n and s[self.i] != "\n":
                        self.i += 1
                else:
                    break
            if self.i >= n:
                return (END, "", f32(0))
            c = s[self.i]
            if c in "[]":
                self.i += 1
                return (OPEN if c == "[" else CLOSE, c, f32(0))
            if c == '"':
                self.i += 1
                j = self.i
                while self.i < n and s[self.i] != '"':
                    self.i += 1
                t = s[j:self.i]
                if self.i < n:
                    self.i += 1
                return (STR, t, f32(0))
            if c in "-+." or c.isdigit():
                j = self.i
                if c in "-+":
                    self.i += 1
                while self.i < n and (s[self.i].isdigit() or s[self.i] in ".eE-+"):
                    self.i += 1
                t = s[j:self.i]
                return (NUM, t, _strtof(t))
            if c.isalpha() or c == "_":
                j = self.i
                while self.i < n and (s[self.i].isalnum() or s[self.i] == "_"):
                    self.i += 1
                return (WORD, s[j:self.i], f32(0))
            self.i += 1

    def peek(self):
        k = self.i
        t = self.next()
        self.i = k
        return t


def read_param(lx):
    """parse_param (pbrt_parser.cpp:104-160) -> (type, name, floats, ints, strings)."""
    kind, text, _ = lx.next()
    p = {"type": "", "name": "", "floats": [], "ints": [], "strings": []}
    if kind != STR:
        return p
    if " " in text:
        p["type"], p["name"] = text.split(" ", 1)
    else:
        p["name"] = text
    integral = p["type"] in ("integer", "bool")
    vk, vt, vn = lx.peek()
    if vk == STR and p["type"] in ("string", "texture"):
        lx.next()
        p["strings"].append(vt)
        return p
    if vk == NUM and (p["type"] == "float" or integral):
        lx.next()
        (p["ints"] if integral else p["floats"]).append(int(vn) if integral else vn)
        return p
    if vk != OPEN:
        return p
    lx.next()
    while True:
        k, t, v = lx.peek()
        if k == CLOSE:
            lx.next()
            break
        if k == END:
            break
        lx.next()
        if k == NUM:
            (p["ints"] if integral else p["floats"]).append(int(v) if integral else v)
        elif k == STR:
            p["strings"].append(t)
    return p


def params(lx):
    while lx.peek()[0] == STR:
        yield read_param(lx)


SPECTRA = {"metal-Ag-eta": (0.155, 0.116, 0.138), "metal-Ag-k": (4.828, 3.122, 2.146),
           "metal-Au-eta": (0.143, 0.374, 1.442), "metal-Au-k": (3.983, 2.387, 1.603),
           "metal-Cu-eta": (0.200, 0.924, 1.102), "metal-Cu-k": (3.912, 2.452, 2.142),
           "metal-Al-eta": (1.345, 0.965, 0.612), "metal-Al-k": (7.474, 6.400, 5.275),
           "metal-Fe-eta": (2.912, 2.950, 2.584), "metal-Fe-k": (3.070, 2.930, 2.771)}


def spectrum(name):
    return tuple(f32(x) for x in SPECTRA.get(name, (0.0, 0.0, 0.0)))


def identity16():
    return [f32(1) if i % 5 == 0 else f32(0) for i in range(16)]


def mul16(a, b):
    r = [f32(0)] * 16
    for i in range(4):
        for j in range(4):
            s = f32(0)
            for k in range(4):
                s = f32(s + f32(a[i * 4 + k] * b[k * 4 + j]))
            r[i * 4 + j] = s
    return r


def splice(path, root, depth=0):
    """preprocess_includes (pbrt_parser.cpp:201-231)."""
    if depth > 10 or not os.path.exists(path):
        return ""
    out = []
    with open(path, newline="") as fh:  # lines as std::getline splits them
        lines = fh.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for line in lines:
        t = line.lstrip(" \t")
        if len(t) > 5 and (t.startswith("Include") or t.startswith("Import")):
            q1, q2 = t.find('"'), t.rfind('"')
            if q1 >= 0 and q2 > q1:
                out.append(splice(root + t[q1 + 1:q2], root, depth + 1))
                continue
        out.append(line + "\n")
    return "".join(out)


def read_ply(path):
    """load_ply (ply_reader.cpp:64-251) -> (positions, normals, uvs, indices) or None."""
    if not os.path.exists(path):
        return None
    data = open(path, "rb").read()
    lines, pos = [], 0
    first_end = data.index(b"\n") + 1
    if b"ply" not in data[:first_end]:
        return None
    pos = first_end
    fmt, elems = "binary_little_endian", []
    while pos < len(data):
        e = data.index(b"\n", pos)
        line = data[pos:e].decode(errors="replace")
        pos = e + 1
        if not line or line[0] == "\r":
            continue
        line = line.rstrip("\r")
        w = line.split()
        if not w:
            continue
        if w[0] == "end_header":
            break
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elems.append({"name": w[1], "count": int(w[2]), "props": []})
        elif w[0] == "property" and elems:
            if w[1] == "list":
                elems[-1]["props"].append({"list": True, "ct": w[2], "vt": w[3], "name": w[4]})
            else:
                elems[-1]["props"].append({"list": False, "type": w[1], "name": w[2]})
    if fmt == "binary_big_endian":
        return None
    # the last element of each name, as the reference's search leaves it
    ve = next((e for e in reversed(elems) if e["name"] == "vertex"), None)
    fe = next((e for e in reversed(elems) if e["name"] == "face"), None)
    if ve is None or fe is None:
        return None
    names = [p["name"] for p in ve["props"]]

    def find(*keys):
        for i, n in enumerate(names):
            if n in keys:
                return i
        return -1
    ix, iy, iz = find("x"), find("y"), find("z")
    inx, iny, inz = find("nx"), find("ny"), find("nz")
    iu, iv = find("u", "s", "texture_u"), find("v", "t", "texture_v")
    has_n, has_uv = min(inx, iny, inz) >= 0, min(iu, iv) >= 0
    nv = ve["count"]
    P, N, UV, I = [f32(0)] * (nv * 3), ([f32(0)] * (nv * 3) if has_n else []), ([f32(0)] * (nv * 2) if has_uv else []), []
    size = {"double": 8, "float64": 8, "int64": 8, "uint64": 8, "uchar": 1, "uint8": 1, "char": 1, "int8": 1,
            "short": 2, "int16": 2, "ushort": 2, "uint16": 2}
    ints = {"uchar": "<B", "uint8": "<B", "short": "<h", "int16": "<h", "ushort": "<H", "uint16": "<H",
            "uint": "<I", "uint32": "<I"}

    def fan(f):
        for j in range(1, len(f) - 1):
            I.extend([f[0], f[j], f[j + 1]])
    if fmt == "ascii":
        toks = data[pos:].split()
        k = 0
        for v in range(nv):
            vals = [f32(0)] * 16
            for p in range(len(names)):
                x = _strtof(toks[k].decode())
                k += 1
                if p < 16:
                    vals[p] = x
            _store(vals, v, P, N, UV, ix, iy, iz, inx, iny, inz, iu, iv, has_n, has_uv)
        for _ in range(fe["count"]):
            n = int(toks[k])
            k += 1
            fan([int(t) for t in toks[k:k + n]])
            k += n
    else:
        for v in range(nv):
            vals = [f32(0)] * 16
            for p, pr in enumerate(ve["props"]):
                if pr["type"] in ("float", "float32"):
                    x = f32(struct.unpack_from("<f", data, pos)[0])
                    pos += 4
                else:
                    x = f32(0)
                    pos += size.get(pr["type"], 4)
                if p < 16:
                    vals[p] = x
            _store(vals, v, P, N, UV, ix, iy, iz, inx, iny, inz, iu, iv, has_n, has_uv)

        def rint(t):
            nonlocal pos
            fm = ints.get(t, "<i")
            v = struct.unpack_from(fm, data, pos)[0]
            pos += struct.calcsize(fm)
            return int(np.int32(np.uint32(v))) if fm == "<I" else v
        for _ in range(fe["count"]):
            for pr in fe["props"]:
                if pr["list"] and pr["name"] in ("vertex_indices", "vertex_index"):
                    n = rint(pr["ct"])
                    fan([rint(pr["vt"]) for _ in range(n)])
                elif pr["list"]:
                    n = rint(pr["ct"])
                    pos += n * size.get(pr["vt"], 4)
                else:
                    pos += size.get(pr["type"], 4)
    if not P or not I:
        return None
    return P, N, UV, I


def _store(vals, v, P, N, UV, ix, iy, iz, inx, iny, inz, iu, iv, has_n, has_uv):
    if ix >= 0:
        P[v * 3] = vals[ix]
    if iy >= 0:
        P[v * 3 + 1] = vals[iy]
    if iz >= 0:
        P[v * 3 + 2] = vals[iz]
    if has_n:
        N[v * 3:v * 3 + 3] = [vals[inx], vals[iny], vals[inz]]
    if has_uv:
        UV[v * 2:v * 2 + 2] = [vals[iu], vals[iv]]


def read_pfm(path):
    """load_pfm (image_loader.cpp:19-77) -> (w, h, rows-from-top RGB float32) or None."""
    if not os.path.exists(path):
        return None
    data = open(path, "rb").read()
    lines = data.split(b"\n", 3)
    magic = lines[0].decode()
    if magic not in ("PF", "Pf"):
        return None
    ch = 3 if magic == "PF" else 1
    w, h = (int(x) for x in lines[1].split()[:2])
    mag = f32(abs(_strtof(lines[2].split()[0].decode())))
    if mag == 0:
        mag = f32(1)
    body = np.frombuffer(lines[3][:w * h * ch * 4], dtype="<f4").reshape(h, w, ch)[::-1]
    rgb = np.repeat(body, 3, axis=2) if ch == 1 else body
    return w, h, (rgb.astype(np.float32) * mag).astype(np.float32)


def parse(path, use_instancing=False):
    root = os.path.dirname(path) + "/" if "/" in path else ""
    lx = Lexer(splice(path, root))
    f = {"width": 512, "height": 512, "spp": 64, "max_depth": 8,
         "cam": {"fov": f32(45), "lensradius": f32(0), "focaldistance": f32(0), "transform": identity16(),
                 "from": (f32(0),) * 3, "at": (f32(0), f32(0), f32(1)), "up": (f32(0), f32(1), f32(0)),
                 "lookat": False},
         "textures": {}, "materials": {}, "meshes": [], "objects": {}, "instances": [],
         "env": False, "env_file": "", "env_transform": identity16(), "env_scale": f32(1)}
    gs = {"material": "", "ctm": identity16(), "L": (f32(0),) * 3, "area": False}
    stack, open_obj, per_obj, tris = [], "", {}, 0

    def new_mesh():
        return {"P": [], "N": [], "UV": [], "I": [], "material": gs["material"], "alpha": "",
                "emission": gs["L"] if gs["area"] else (f32(0),) * 3, "emissive": gs["area"],
                "transform": list(gs["ctm"])}

    def add(m):
        nonlocal tris
        if open_obj:
            f["objects"].setdefault(open_obj, []).append(m)
        else:
            tris += len(m["I"]) // 3
            f["meshes"].append(m)
    while True:
        k, d, _ = lx.next()
        if k == END:
            break
        if k != WORD:
            continue
        if d in ("Integrator", "Sampler", "Film"):
            lx.next()
            keys = {"Integrator": {"maxdepth": "max_depth"}, "Sampler": {"pixelsamples": "spp"},
                    "Film": {"xresolution": "width", "yresolution": "height"}}[d]
            for p in params(lx):
                if p["name"] in keys and p["ints"]:
                    f[keys[p["name"]]] = p["ints"][0]
        elif d == "PixelFilter":
            lx.next()
            list(params(lx))
        elif d == "Camera":
            lx.next()
            f["cam"]["transform"] = list(gs["ctm"])
            for p in params(lx):
                if p["floats"] and p["name"] in ("fov", "lensradius", "focaldistance"):
                    f["cam"][p["name"]] = p["floats"][0]
        elif d == "Transform":
            if lx.next()[0] == OPEN:
                gs["ctm"] = [lx.next()[2] for _ in range(16)]
                lx.next()
        elif d == "LookAt":
            v = [lx.next()[2] for _ in range(9)]
            f["cam"].update({"from": tuple(v[0:3]), "at": tuple(v[3:6]), "up": tuple(v[6:9]), "lookat": True})
        elif d == "Scale":
            sx, sy, sz = (lx.next()[2] for _ in range(3))
            gs["ctm"] = mul16(gs["ctm"], [f32(x) for x in (sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, sz, 0, 0, 0, 0, 1)])
        elif d == "Rotate":
            ang = lx.next()[2]
            ax, ay, az = (lx.next()[2] for _ in range(3))
            ln = _sqrt(f32(f32(f32(ax * ax) + f32(ay * ay)) + f32(az * az)))
            if ln > 0:
                ax, ay, az = f32(ax / ln), f32(ay / ln), f32(az / ln)
            rad = f32(f32(ang * f32(3.14159265)) / f32(180))
            c, s = _cos(rad), _sin(rad)
            u = f32(1 - c)
            m = lambda a, b: f32(f32(u * a) * b)  # noqa: E731
            r = [f32(m(ax, ax) + c), f32(m(ax, ay) - f32(s * az)), f32(m(ax, az) + f32(s * ay)), 0,
                 f32(m(ax, ay) + f32(s * az)), f32(m(ay, ay) + c), f32(m(ay, az) - f32(s * ax)), 0,
                 f32(m(ax, az) - f32(s * ay)), f32(m(ay, az) + f32(s * ax)), f32(m(az, az) + c), 0,
                 0, 0, 0, 1]
            gs["ctm"] = mul16(gs["ctm"], [f32(x) for x in r])
        elif d == "ConcatTransform":
            if lx.next()[0] == OPEN:
                m = [lx.next()[2] for _ in range(16)]
                lx.next()
                gs["ctm"] = mul16(m, gs["ctm"])
        elif d in ("Identity", "WorldBegin"):
            gs["ctm"] = identity16()
        elif d == "AttributeBegin":
            stack.append(dict(gs, ctm=list(gs["ctm"])))
        elif d == "AttributeEnd":
            if stack:
                gs = stack.pop()
        elif d == "MakeNamedMaterial":
            name = lx.next()[1]
            m = {"type": "", "refl": (f32(0.5),) * 3, "eta": (f32(0),) * 3, "k": (f32(0),) * 3,
                 "rough": f32(0), "urough": f32(-1), "vrough": f32(-1), "eta_s": f32(1.5), "refl_tex": ""}
            for p in params(lx):
                n, fl, st = p["name"], p["floats"], p["strings"]
                if n == "type" and st:
                    m["type"] = st[0]
                elif n == "reflectance" and p["type"] == "texture" and st:
                    m["refl_tex"] = st[0]
                elif n == "reflectance" and len(fl) >= 3:
                    m["refl"] = tuple(fl[:3])
                elif n == "roughness" and fl:
                    m["rough"] = fl[0]
                elif n == "eta":
                    if st:
                        m["eta"] = spectrum(st[0])
                    elif len(fl) >= 3:
                        m["eta"] = tuple(fl[:3])
                    elif len(fl) == 1:
                        m["eta_s"] = fl[0]
                elif n == "k":
                    if st:
                        m["k"] = spectrum(st[0])
                    elif len(fl) >= 3:
                        m["k"] = tuple(fl[:3])
                elif n == "uroughness" and fl:
                    m["urough"] = fl[0]
                elif n == "vroughness" and fl:
                    m["vrough"] = fl[0]
            f["materials"][name] = m
        elif d == "NamedMaterial":
            gs["material"] = lx.next()[1]
        elif d == "AreaLightSource":
            lx.next()
            gs["area"] = True
            for p in params(lx):
                if p["name"] == "L" and len(p["floats"]) >= 3:
                    gs["L"] = tuple(p["floats"][:3])
        elif d == "Shape":
            typ = lx.next()[1]
            if typ == "trianglemesh":
                m = new_mesh()
                for p in params(lx):
                    if p["name"] in ("P", "p") and p["type"] == "point3":
                        m["P"] = p["floats"]
                    elif p["name"] == "indices" and p["type"] == "integer":
                        m["I"] = p["ints"]
                    elif p["name"] == "N" and p["type"] == "normal":
                        m["N"] = p["floats"]
                    elif p["name"] == "uv" and p["type"] == "point2":
                        m["UV"] = p["floats"]
                if m["P"] and m["I"]:
                    add(m)
            elif typ == "plymesh":
                m = new_mesh()
                fn = ""
                for p in params(lx):
                    if p["name"] == "filename" and p["strings"]:
                        fn = p["strings"][0]
                    elif p["name"] == "alpha" and p["type"] == "texture" and p["strings"]:
                        m["alpha"] = p["strings"][0]
                if fn:
                    r = read_ply(root + fn)
                    if r is not None:
                        m["P"], m["N"], m["UV"], m["I"] = r
                        add(m)
            elif typ == "sphere":
                m = new_mesh()
                radius = f32(1)
                for p in params(lx):
                    if p["name"] == "radius" and p["floats"]:
                        radius = p["floats"][0]
                for j in range(17):
                    th = f32(f32(f32(j) / f32(16)) * f32(3.14159265))
                    st_, ct = _sin(th), _cos(th)
                    for i in range(33):
                        ph = f32(f32(f32(i) / f32(32)) * f32(6.28318530))
                        cp, sp = _cos(ph), _sin(ph)
                        m["P"] += [f32(f32(st_ * cp) * radius), f32(ct * radius), f32(f32(st_ * sp) * radius)]
                        m["N"] += [f32(st_ * cp), ct, f32(st_ * sp)]
                for j in range(16):
                    for i in range(32):
                        a = j * 33 + i
                        b, c = a + 1, a + 33
                        m["I"] += [a, c, b, b, c, c + 1]
                add(m)
            else:
                list(params(lx))
        elif d == "Texture":
            name = lx.next()[1]
            vt, ct = lx.next()[1], lx.next()[1]
            t = {"class": ct, "file": "", "img": None}
            for p in params(lx):
                if p["name"] == "filename" and p["strings"]:
                    t["file"] = p["strings"][0]
            f["textures"][name] = t
        elif d == "LightSource":
            typ = lx.next()[1]
            if typ == "infinite":
                f["env"] = True
                f["env_transform"] = list(gs["ctm"])
                for p in params(lx):
                    if p["name"] == "filename" and p["strings"]:
                        f["env_file"] = p["strings"][0]
                    elif p["name"] in ("scale", "L") and len(p["floats"]) >= 3:
                        a, b, c = p["floats"][:3]
                        f["env_scale"] = f32(f32(f32(a + b) + c) / f32(3))
            else:
                list(params(lx))
        elif d == "ObjectBegin":
            open_obj = lx.next()[1]
        elif d == "ObjectEnd":
            open_obj = ""
        elif d == "ObjectInstance":
            name = lx.next()[1]
            if name not in f["objects"]:
                continue
            f["instances"].append(name)
            if use_instancing or tris >= 500000000:
                continue
            if per_obj.get(name, 0) >= 500:
                continue
            ot = sum(len(m["I"]) // 3 for m in f["objects"][name])
            if tris + ot > 500000000:
                continue
            for m in f["objects"][name]:
                f["meshes"].append(dict(m, transform=list(gs["ctm"])))
            tris += ot
            per_obj[name] = per_obj.get(name, 0) + 1
        elif d in ("Material", "MakeNamedMedium"):
            lx.next()
            list(params(lx))
        elif d in ("ReverseOrientation", "TransformBegin", "TransformEnd", "MediumInterface"):
            list(params(lx))
    f["root"] = root
    return f


# ---- conversion (pbrt_to_photon.cpp:58-461) ----------------------------------------
def _cols(cm):
    return [[f32(cm[c * 4 + r]) for c in range(4)] for r in range(4)]


def _point(m, p):
    x, y, z = p
    o = [f32(f32(f32(f32(m[r][0] * x) + f32(m[r][1] * y)) + f32(m[r][2] * z)) + m[r][3]) for r in range(4)]
    if o[3] != 0 and o[3] != 1:
        iw = f32(f32(1) / o[3])
        return (f32(o[0] * iw), f32(o[1] * iw), f32(o[2] * iw))
    return (o[0], o[1], o[2])


def _dir(m, d):
    x, y, z = d
    return tuple(f32(f32(f32(m[r][0] * x) + f32(m[r][1] * y)) + f32(m[r][2] * z)) for r in range(3))


def _sub(a, b):
    return tuple(f32(x - y) for x, y in zip(a, b))


def _cross(a, b):
    return (f32(f32(a[1] * b[2]) - f32(a[2] * b[1])), f32(f32(a[2] * b[0]) - f32(a[0] * b[2])),
            f32(f32(a[0] * b[1]) - f32(a[1] * b[0])))


def _len(a):
    return _sqrt(f32(f32(f32(a[0] * a[0]) + f32(a[1] * a[1])) + f32(a[2] * a[2])))


def _norm(a):
    ln = _len(a)
    if ln > 0:
        s = f32(f32(1) / ln)
        return tuple(f32(x * s) for x in a)
    return (f32(0),) * 3


def _inverse(m):
    a = m
    det = f32(f32(f32(a[0][0] * f32(f32(a[1][1] * a[2][2]) - f32(a[1][2] * a[2][1])))
                  - f32(a[0][1] * f32(f32(a[1][0] * a[2][2]) - f32(a[1][2] * a[2][0]))))
              + f32(a[0][2] * f32(f32(a[1][0] * a[2][1]) - f32(a[1][1] * a[2][0]))))
    r = [[f32(0)] * 4 for _ in range(4)]
    if det == 0:
        return r
    r[3][3] = f32(1)
    i = f32(f32(1) / det)
    d = lambda p, q, s, t: f32(f32(f32(p * q) - f32(s * t)) * i)  # noqa: E731
    r[0][0] = d(a[1][1], a[2][2], a[1][2], a[2][1])
    r[0][1] = d(a[0][2], a[2][1], a[0][1], a[2][2])
    r[0][2] = d(a[0][1], a[1][2], a[0][2], a[1][1])
    r[1][0] = d(a[1][2], a[2][0], a[1][0], a[2][2])
    r[1][1] = d(a[0][0], a[2][2], a[0][2], a[2][0])
    r[1][2] = d(a[0][2], a[1][0], a[0][0], a[1][2])
    r[2][0] = d(a[1][0], a[2][1], a[1][1], a[2][0])
    r[2][1] = d(a[0][1], a[2][0], a[0][0], a[2][1])
    r[2][2] = d(a[0][0], a[1][1], a[0][1], a[1][0])
    t = (a[0][3], a[1][3], a[2][3])
    for k in range(3):
        r[k][3] = f32(-f32(f32(f32(r[k][0] * t[0]) + f32(r[k][1] * t[1])) + f32(r[k][2] * t[2])))
    return r


def convert(f):
    """-> dict of the converted arrays, materials (abi.Material), lights,
    camera (abi.Camera) and settings."""
    import oracle
    from photon_amd import abi
    root = f["root"]
    texs = []
    tex_id = {}
    for name in sorted(f["textures"]):
        t = f["textures"][name]
        if t["class"] == "imagemap" and t["file"]:
            t["img"] = read_pfm(root + t["file"]) if t["file"].lower().endswith(".pfm") else None
        if t["img"] is not None:
            tex_id[name] = len(texs)
            texs.append(t["img"])
    mats, mat_id = [], {}

    def rough(p):
        if p["urough"] >= 0:
            v = p["vrough"] if p["vrough"] >= 0 else p["urough"]
            return f32(f32(p["urough"] + v) * f32(0.5))
        return p["rough"]
    for name in sorted(f["materials"]):
        p = f["materials"][name]
        mat_id[name] = len(mats)
        m = abi.default_material()
        if p["refl_tex"] in tex_id:
            m.base_color_tex = tex_id[p["refl_tex"]]
        typ = p["type"]
        if typ == "diffuse":
            m.base_color[:] = p["refl"]
            m.roughness, m.metallic = 1.0, 0.0
        elif typ == "conductor":
            def f0(eta, k):
                num = f32(f32(f32(eta - 1) * f32(eta - 1)) + f32(k * k))
                den = f32(f32(f32(eta + 1) * f32(eta + 1)) + f32(k * k))
                return f32(num / den) if den > 0 else f32(0)
            m.base_color[:] = tuple(float(f0(e, k)) for e, k in zip(p["eta"], p["k"]))
            m.metallic = 1.0
            m.roughness = rough(p)
        elif typ == "dielectric":
            m.base_color[:] = (1.0, 1.0, 1.0)
            m.ior, m.transmission, m.roughness = p["eta_s"], 1.0, rough(p)
        elif typ == "coateddiffuse":
            m.base_color[:] = p["refl"]
            r = rough(p)
            m.metallic, m.roughness, m.clearcoat, m.clearcoat_roughness = 0.0, r, 1.0, f32(r * f32(0.5))
        elif typ == "diffusetransmission":
            m.base_color[:] = p["refl"]
            m.roughness, m.metallic, m.transmission = 1.0, 0.0, 0.5
        else:
            m.base_color[:] = p["refl"]
            m.roughness = p["rough"]
        mats.append(m)
    if not mats:
        m = abi.default_material()
        m.base_color[:] = (0.5, 0.5, 0.5)
        m.roughness = 1.0
        mats.append(m)
    P, N, UV, I, MID, EID, EA = [], [], [], [], [], [], []
    for mesh in f["meshes"]:
        xf = _cols(mesh["transform"])
        mid = mat_id.get(mesh["material"], 0)
        pos = mesh["P"]
        nv = len(pos) // 3
        has_uv, has_n = len(mesh["UV"]) >= nv * 2, len(mesh["N"]) >= len(pos)
        if mesh["alpha"] in tex_id:
            am = abi.Material.from_buffer_copy(mats[mid])
            am.alpha_tex = tex_id[mesh["alpha"]]
            mid = len(mats)
            mats.append(am)
        if mesh["emissive"]:
            em = abi.default_material()
            em.base_color[:] = (0.0, 0.0, 0.0)
            em.emission[:] = mesh["emission"]
            em.emission_strength, em.roughness = 1.0, 1.0
            mid = len(mats)
            mats.append(em)
        idx = mesh["I"]
        for t in range(len(idx) // 3):
            ii = idx[t * 3:t * 3 + 3]
            if min(ii) < 0 or max(ii) >= nv:
                continue
            p = [_point(xf, tuple(f32(x) for x in pos[i * 3:i * 3 + 3])) for i in ii]
            base = len(P)
            P += p
            I.append((base, base + 1, base + 2))
            if has_n:
                N += [_norm(_dir(xf, tuple(f32(x) for x in mesh["N"][i * 3:i * 3 + 3]))) for i in ii]
            else:
                n = _norm(_cross(_sub(p[1], p[0]), _sub(p[2], p[0])))
                N += [n, n, n]
            UV += [tuple(mesh["UV"][i * 2:i * 2 + 2]) for i in ii] if has_uv else [(f32(0), f32(0))] * 3
            if mesh["emissive"]:
                EID.append(len(MID))
                EA.append(f32(f32(0.5) * _len(_cross(_sub(p[1], p[0]), _sub(p[2], p[0])))))
            MID.append(mid)
    lights = []
    if EID:
        total = f32(0)
        for a in EA:
            total = f32(total + a)
        for mesh in f["meshes"]:
            if not mesh["emissive"]:
                continue
            xf = _cols(mesh["transform"])
            pos, idx = mesh["P"], mesh["I"]
            nv = len(pos) // 3
            ok = lambda k: k < len(idx) and 0 <= idx[k] < nv  # noqa: E731
            if not (ok(0) and ok(1)):
                continue
            v = lambda i: _point(xf, tuple(f32(x) for x in pos[i * 3:i * 3 + 3]))  # noqa: E731
            p0, p1 = v(idx[0]), v(idx[1])
            p3 = v(idx[5]) if len(idx) >= 6 and ok(5) else p0
            l = abi.Light()
            l.spot_angle, l.spot_falloff = 0.5, 0.1
            l.type = abi.PHX_LIGHT_AREA
            e1, e2 = _sub(p1, p0), _sub(p3, p0)
            cr = _cross(e1, e2)
            cl = _len(cr)
            l.direction[:] = tuple(f32(x * f32(f32(1) / cl)) for x in cr) if cl > 0 else (0.0, -1.0, 0.0)
            l.position[:], l.color[:] = p0, mesh["emission"]
            l.intensity, l.area = 1.0, total
            l.mesh_prim_begin, l.mesh_prim_count = EID[0], len(idx) // 3
            l.edge1[:], l.edge2[:] = e1, e2
            lights.append(l)
    env = None
    w2t = np.eye(3, dtype=np.float32).reshape(-1)
    if f["env"] and f["env_file"] and f["env_file"].lower().endswith(".pfm"):
        img = read_pfm(root + f["env_file"])
        if img is not None:
            env = (img[2] * f["env_scale"]).astype(np.float32)
            inv = _inverse(_cols(f["env_transform"]))
            w2t = np.array([inv[r][c] for r in range(3) for c in range(3)], np.float32)
    cam = f["cam"]
    if cam["lookat"]:
        pos = cam["from"]
        d = _norm(_sub(cam["at"], pos))
        up = _norm(cam["up"])
    else:
        c2w = _inverse(_cols(cam["transform"]))
        pos = _point(c2w, (f32(0),) * 3)
        d = _norm(_dir(c2w, (f32(0), f32(0), f32(1))))
        up = _norm(_dir(c2w, (f32(0), f32(1), f32(0))))
    aspect = f32(f32(f["width"]) / f32(f["height"]))
    at = tuple(f32(a + b) for a, b in zip(pos, d))
    camera = oracle.camera_perspective(pos, at, up, float(cam["fov"]), float(aspect), 0.0, 1.0)
    if cam["lensradius"] > 0:
        camera.lens_radius = cam["lensradius"]
        camera.focus_dist = cam["focaldistance"] if cam["focaldistance"] > 0 else _len(d)
    if not I:
        P, N, UV, I, MID = [(f32(0),) * 3] * 3, [(f32(0), f32(1), f32(0))] * 3, [(f32(0), f32(0))] * 3, [(0, 1, 2)], [0]
    tex_pixels = np.concatenate([t[2].reshape(-1, 3) for t in texs]) if texs else np.zeros((0, 3), np.float32)
    infos, off = [], 0
    for t in texs:
        infos.append((off, t[0], t[1]))
        off += t[0] * t[1]
    return {"positions": np.array(P, np.float32).reshape(-1, 3), "indices": np.array(I, np.uint32).reshape(-1, 3),
            "normals": np.array(N, np.float32).reshape(-1, 3), "texcoords": np.array(UV, np.float32).reshape(-1, 2),
            "material_ids": np.array(MID, np.uint32), "emissive_ids": np.array(EID, np.uint32),
            "emissive_areas": np.array(EA, np.float32), "materials": mats, "lights": lights,
            "texture_pixels": tex_pixels, "texture_infos": infos,
            "env_pixels": np.zeros(0, np.float32) if env is None else env.reshape(-1),
            "env_size": (0, 0) if env is None else (env.shape[1], env.shape[0]), "env_world_to_tex": w2t,
            "camera": camera,
            "info": (max(f["width"], 1), max(f["height"], 1), max(f["spp"], 1), max(f["max_depth"], 1))}


def scene_of(conv, name="pbrt"):
    """The converted arrays as a photon_amd.scenes.Scene (the oracle renders it)."""
    from photon_amd import scenes
    s = scenes.Scene(conv["positions"], conv["indices"], conv["materials"], material_ids=conv["material_ids"],
                     normals=conv["normals"], texcoords=conv["texcoords"])
    s.lights = list(conv["lights"])
    if conv["emissive_ids"].size:
        s.emissive_ids, s.emissive_areas = conv["emissive_ids"], conv["emissive_areas"]
    if conv["texture_infos"]:
        s.texture_pixels = conv["texture_pixels"]
        s.texture_infos = conv["texture_infos"]
    if conv["env_size"][0]:
        w, h = conv["env_size"]
        s.env_pixels = conv["env_pixels"].reshape(h, w, 3)
        s.env_world_to_tex = conv["env_world_to_tex"].reshape(3, 3)
    s.name = name
    return s
