"""SPIDER file formats at the operation boundary: images, volumes and stacks
(docs/image_doc.html) and document files (docs/docfile.html).

Only what the `BP 3F` / `BP 32F` signature needs: read simple 2D images,
image stacks, indexed stacks and templates (`jnk***`, `stk@***`), write simple volumes and
stacks, read/write document files, and the raw Fourier partial volumes that
`BP 3F P` writes and `BP 3F M` reads (bp3f.f:191-248, 646-685).
Both byte orders are read; files are written in the machine's order.
"""
from __future__ import annotations

import datetime
import os
import re

import numpy as np

# header word positions (1-based, docs/image_doc.html)
_NZ, _NY, _IREC, _IFORM, _IMAMI, _FMAX, _FMIN, _AV, _SIG = 1, 2, 3, 5, 6, 7, 8, 9, 10
_NX, _LABREC, _LABBYT, _LENBYT, _ISTACK, _MAXIM, _IMGNUM = 12, 13, 22, 23, 24, 26, 27


class SpiderFormatError(ValueError):
    pass


def _header_len(nx: int) -> tuple[int, int]:
    """LABREC and LABBYT for a record length of NX words (image_doc.html)."""
    lenbyt = nx * 4
    labrec = 1024 // lenbyt
    if 1024 % lenbyt != 0:
        labrec += 1
    return labrec, labrec * lenbyt


def read_header(path: str) -> dict:
    """Header of a SPIDER file; detects the byte order."""
    raw = np.fromfile(path, dtype=np.uint8, count=1024)
    if raw.size < 1024 and raw.size < 4 * 24:
        raise SpiderFormatError(f"{path}: too short for a SPIDER header")
    for bo in ("<", ">"):
        h = np.frombuffer(raw[: (raw.size // 4) * 4].tobytes(), dtype=bo + "f4")
        nx, ny, nz, iform = h[_NX - 1], h[_NY - 1], h[_NZ - 1], h[_IFORM - 1]
        labbyt, lenbyt = h[_LABBYT - 1], h[_LENBYT - 1]
        if (nx >= 1 and ny >= 1 and nx == int(nx) and ny == int(ny) and lenbyt == 4 * nx
                and labbyt > 0 and labbyt == int(labbyt) and int(iform) in (1, 3, -11, -12, -21, -22)):
            return {
                "byteorder": bo, "nx": int(nx), "ny": int(ny), "nz": max(1, int(nz)), "iform": int(iform),
                "labbyt": int(labbyt), "lenbyt": int(lenbyt), "istack": int(h[_ISTACK - 1]),
                "maxim": int(h[_MAXIM - 1]), "imgnum": int(h[_IMGNUM - 1]),
            }
    raise SpiderFormatError(f"{path}: not a SPIDER image file")


def _indexed_slot(path: str, h: dict, index: int | None) -> int:
    """Storage slot (1-based) of image `index` in an indexed stack
    (docs/image_doc.html): ISTACK = -MAXINDX in the overall header; after
    it, INDXREC = ceil(MAXINDX / NX) records of NX 32-bit integers give each
    image number's slot, 0 where the image does not exist (LUNGETLAB,
    LUNREDINDX, lunsethdr.f)."""
    maxindx = -h["istack"]
    if index is None:
        raise SpiderFormatError(f"{path}: is a stack, give an image number")
    if not 1 <= index <= maxindx:
        raise SpiderFormatError(f"{path}: image number {index} exceeds index limit {maxindx}")
    rec, loc = divmod(index - 1, h["nx"])
    idx = np.fromfile(path, dtype=np.dtype(h["byteorder"] + "i4"), count=1,
                      offset=h["labbyt"] + rec * h["lenbyt"] + loc * 4)
    if idx.size != 1 or idx[0] <= 0:
        raise SpiderFormatError(f"{path}: no image {index} in the stack")
    return int(idx[0])


def read_image(path: str, index: int | None = None) -> np.ndarray:
    """A simple image/volume, or image `index` (1-based) of a stack or an
    indexed stack: float32 array (NY, NX) for 2D or (NZ, NY, NX) for 3D."""
    h = read_header(path)
    nx, ny, nz, lab = h["nx"], h["ny"], h["nz"], h["labbyt"]
    n = nx * ny * nz
    dt = np.dtype(h["byteorder"] + "f4")
    if h["istack"] < 0:
        slot = _indexed_slot(path, h, index)
        indxrec = (-h["istack"] + nx - 1) // nx                 # INDXREC = ceil(MAXINDX / NX) records
        off = lab + indxrec * h["lenbyt"] + (slot - 1) * (lab + n * 4) + lab
    elif h["istack"] > 0:
        if index is None:
            raise SpiderFormatError(f"{path}: is a stack, give an image number")
        if not 1 <= index <= h["maxim"]:
            raise SpiderFormatError(f"{path}: no image {index} (stack holds {h['maxim']})")
        off = lab + (index - 1) * (lab + n * 4) + lab
    else:
        if index not in (None, 1):
            raise SpiderFormatError(f"{path}: not a stack")
        off = lab
    a = np.fromfile(path, dtype=dt, count=n, offset=off)
    if a.size != n:
        raise SpiderFormatError(f"{path}: truncated")
    a = a.astype(np.float32).reshape(nz, ny, nx)
    return a[0] if nz == 1 else a


def _make_header(nx: int, ny: int, nz: int, iform: int, data: np.ndarray | None, istack: int = 0,
                 maxim: int = 0, imgnum: int = 0, irec: int | None = None) -> np.ndarray:
    labrec, labbyt = _header_len(nx)
    h = np.zeros(labbyt // 4, np.float32)
    h[_NZ - 1], h[_NY - 1], h[_IFORM - 1], h[_NX - 1] = nz, ny, iform, nx
    h[_IREC - 1] = labrec + ny * nz if irec is None else irec
    h[_LABREC - 1], h[_LABBYT - 1], h[_LENBYT - 1] = labrec, labbyt, nx * 4
    h[_ISTACK - 1], h[_MAXIM - 1], h[_IMGNUM - 1] = istack, maxim, imgnum
    if data is not None and data.size:
        d = data.astype(np.float64)
        h[_IMAMI - 1] = 1
        h[_FMAX - 1], h[_FMIN - 1], h[_AV - 1], h[_SIG - 1] = d.max(), d.min(), d.mean(), d.std()
    if h.size >= 256:
        now = datetime.datetime.now()
        txt = (now.strftime("%d-%b-%Y").upper().ljust(12) + now.strftime("%H:%M:%S").ljust(8)).ljust(180)
        h[211:256] = np.frombuffer(txt.encode("ascii")[:180], dtype=np.float32)
    return h


def write_image(path: str, a) -> None:
    """Simple 2D image (NY, NX) or 3D volume (NZ, NY, NX), IFORM 1 / 3."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    if a.ndim == 2:
        nz, (ny, nx), iform = 1, a.shape, 1
    elif a.ndim == 3:
        (nz, ny, nx), iform = a.shape, 3
    else:
        raise SpiderFormatError("write_image needs a 2D or 3D array")
    h = _make_header(nx, ny, nz, iform, a)
    with open(path, "wb") as fh:
        h.tofile(fh)
        a.tofile(fh)


def write_stack(path: str, imgs) -> None:
    """Image stack (n, NY, NX) with an overall header and one header per image."""
    imgs = np.ascontiguousarray(np.asarray(imgs, np.float32))
    n, ny, nx = imgs.shape
    labrec, _ = _header_len(nx)
    with open(path, "wb") as fh:
        _make_header(nx, ny, 1, 1, None, istack=2, maxim=n, irec=labrec + ny).tofile(fh)
        for i in range(n):
            _make_header(nx, ny, 1, 1, imgs[i], imgnum=i + 1).tofile(fh)
            imgs[i].tofile(fh)


def write_images(template: str, numbers, imgs, ext: str = "") -> None:
    """Write images (n, NY, NX) under a template, one per file number:
    `name***` -> one simple image per number; `stk@***` -> one stack whose
    image k holds the k-th listed number's image (MAXIM = the largest
    number; slots no number names carry IMGNUM 0, unused).  As the
    transformed-images output of `RB 3F` / `RB 32F` (bp32f.f:517-527,
    826-827)."""
    imgs = np.ascontiguousarray(np.asarray(imgs, np.float32))
    names = template_names(template, numbers, ext)
    if len(names) != imgs.shape[0]:
        raise SpiderFormatError(f"{len(names)} file numbers for {imgs.shape[0]} images")
    if "@" not in template:
        for (f, _), a in zip(names, imgs):
            write_image(f, a)
        return
    n, ny, nx = imgs.shape
    maxim = max(k for _, k in names)
    if min(k for _, k in names) < 1:
        raise SpiderFormatError("stack image numbers start at 1")
    slot = {k: i for i, (_, k) in enumerate(names)}
    labrec, _ = _header_len(nx)
    zero = np.zeros((ny, nx), np.float32)
    with open(names[0][0], "wb") as fh:
        _make_header(nx, ny, 1, 1, None, istack=2, maxim=maxim, irec=labrec + ny).tofile(fh)
        for k in range(1, maxim + 1):
            a = imgs[slot[k]] if k in slot else zero
            _make_header(nx, ny, 1, 1, a if k in slot else None, imgnum=k if k in slot else 0).tofile(fh)
            a.tofile(fh)


def _expand_numbers(spec) -> list[int]:
    """File-number list as SPIDER accepts it: "1-12", "1,3,5-9", or a sequence."""
    if isinstance(spec, str):
        out = []
        for part in re.split(r"[,\s]+", spec.strip()):
            if not part:
                continue
            if "-" in part[1:]:
                a, b = part.split("-", 1) if not part.startswith("-") else (part, part)
                out.extend(range(int(a), int(b) + 1))
            else:
                out.append(int(part))
        return out
    return [int(v) for v in spec]


def template_names(template: str, numbers, ext: str = "") -> list[tuple[str, int | None]]:
    """(file, stack index) per number for `name***` or `stk@***` templates."""
    nums = _expand_numbers(numbers)
    out = []
    if "@" in template:
        base, pat = template.split("@", 1)
        f = base if os.path.splitext(base)[1] or not ext else base + ext
        for k in nums:
            out.append((f, k))
        return out
    m = re.search(r"\*+", template)
    if not m:
        raise SpiderFormatError(f"template {template!r} has no *** field")
    for k in nums:
        name = template[: m.start()] + str(k).zfill(m.end() - m.start()) + template[m.end():]
        if ext and not os.path.splitext(name)[1]:
            name += ext
        out.append((name, None))
    return out


def read_images(template: str, numbers, ext: str = "") -> np.ndarray:
    """Image series of a template as float32 (n, NY, NX)."""
    names = template_names(template, numbers, ext)
    imgs = [read_image(f, idx) for f, idx in names]
    shapes = {a.shape for a in imgs}
    if len(shapes) != 1:
        raise SpiderFormatError(f"images differ in size: {sorted(shapes)}")
    return np.stack(imgs).astype(np.float32)


# ------------------------------------------------------------------ document files
def read_doc(path: str) -> dict[int, list[float]]:
    """{key: [registers]} of a SPIDER document file; ';' in column 2 marks a
    comment line (docfile.html).  Reads the fixed-column pre-11.0 form
    (I6,I1,9G12.3) as well as the blank-separated form."""
    out: dict[int, list[float]] = {}
    with open(path) as fh:
        for line in fh:
            s = line.rstrip("\n")
            if not s.strip() or s[1:2] == ";" or s.lstrip().startswith(";"):
                continue
            parts = s.split()
            try:
                key, n = int(parts[0]), int(parts[1])
                vals = [float(v) for v in parts[2:2 + n]]
                if len(vals) != n:
                    raise ValueError
            except ValueError:
                # pre-11.0 fixed columns: key I6, count I1, registers G12.3
                key, n = int(s[0:6]), int(s[6:7])
                vals = [float(s[7 + 12 * i: 19 + 12 * i]) for i in range(n)]
            out[key] = vals
    return out


def write_doc(path: str, data: dict, comment: str = "") -> None:
    """Document file in the blank-separated form with 1PGE12.5 registers."""
    with open(path, "w") as fh:
        now = datetime.datetime.now().strftime("%d-%b-%Y AT %H:%M:%S").upper()
        fh.write(f" ;spider_amd   {now}   {os.path.basename(path)}\n")
        if comment:
            fh.write(f" ; {comment}\n")
        for key in sorted(data):
            vals = list(data[key])
            fh.write(f"{key:6d} {len(vals):2d} " + " ".join(f"{v:12.5G}" for v in vals) + "\n")


def angles_for(doc: dict, numbers) -> np.ndarray:
    """Registers (PSI, THETA, PHI) per image number; raises for a missing key
    like BUILDM1 (bp3f.f:1280-1288)."""
    out = []
    for k in _expand_numbers(numbers):
        if k not in doc or len(doc[k]) < 3:
            raise KeyError(f"MISSING ANGLE FOR IMAGE {k}")
        out.append(doc[k][:3])
    return np.asarray(out, np.float64)


def inplane_for(doc: dict, numbers) -> np.ndarray:
    """Registers 6, 7, 8 (in-plane angle, x shift, y shift) per image number,
    as RB 3F / RB 32F read them (bp32f.f:122-127, 805-837)."""
    out = []
    for k in _expand_numbers(numbers):
        if k not in doc or len(doc[k]) < 8:
            raise KeyError(f"MISSING ANGLE FOR IMAGE {k} (registers 6-8)")
        out.append(doc[k][5:8])
    return np.asarray(out, np.float64)


# ------------------------------------------------------------------ BP 3F P partial volumes
def write_partial(path: str, X, W) -> str:
    """`BP 3F P` output: the Fourier half grid X (N, N, NX+1) complex as an
    IFORM 3 volume of 2 (NX+1) x N x N reals, and W (N, N, NX+1) in a
    companion file named with _W before the '@' or at the end (bp3f.f:204-210)."""
    X = np.ascontiguousarray(np.asarray(X, np.complex64))
    W = np.ascontiguousarray(np.asarray(W, np.float32))
    write_image(path, X.view(np.float32))
    wpath = _weights_name(path)
    write_image(wpath, W)
    return wpath


def read_partial(path: str):
    X = read_image(path)
    W = read_image(_weights_name(path))
    if X.ndim != 3 or X.shape[2] != 2 * W.shape[2] or X.shape[:2] != W.shape[:2]:
        raise SpiderFormatError(f"{path}: not a BP 3F P partial volume pair")
    return np.ascontiguousarray(X).view(np.complex64), W


def _weights_name(path: str) -> str:
    """Name of a partial volume's weights file: `_W` before the '@' of a
    stack name (`stk@3` -> `stk_W@3`), else at the end of the name
    (bp3f.f:204-210, 576-583); a data extension stays at the end of the file
    name (`stk.dat@3` -> `stk_W.dat@3`, `vol.dat` -> `vol_W.dat`)."""
    if "@" in path:
        base, rest = path.split("@", 1)
        root, ext = os.path.splitext(base)
        return root + "_W" + ext + "@" + rest
    root, ext = os.path.splitext(path)
    return root + "_W" + ext
