"""Write the golden fixtures of iio's image names and transformations.

Calls the reference's own readers, ``iio_read_image_float_split`` and
``iio_read_image_float`` of ``oracle/_ref/libsimpois_ref.so`` (built from
the reference sources by ``oracle/Makefile``, plain ``cc -O3``), on seeded
NPY files of several sample types, under names of the forms
``zero:WxH[,pd]``, ``one:WxH[,pd]``, ``constant:V:WxH`` (iio.c:5603-5634),
``TRANS[list]:name`` (iio.c:3800-3884, 5677-5680) and with the
``IIO_TRANS`` environment variable (iio.c:5702-5704).  Inputs, names,
environment and outputs go to ``tests/golden/iio_names.npz``.  Only names
the reference reads without an error are used (its errors exit the
process).

Needs the reference tree, so it runs where that tree is mounted; the tests
only read the fixture file it writes.

    python tests/golden/make_golden_iio.py
"""
import ctypes
import json
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsimpois_ref.so")


def inputs():
    """name -> array as np.save writes it ((h, w) or (h, w, pd))"""
    rng = np.random.default_rng(2024)
    f1 = (rng.standard_normal((5, 7)) * 10).astype(np.float32)
    f1[1, 2] = np.nan
    f3 = (rng.standard_normal((4, 6, 3)) * 10).astype(np.float32)
    return {
        "f1": f1,                                                  # float, 1 channel
        "f3": f3,                                                  # float, 3 channels
        "u1c3": rng.integers(0, 256, (6, 5, 3)).astype(np.uint8),  # 8-bit colour
        "u1c4": rng.integers(0, 256, (3, 4, 4)).astype(np.uint8),  # 8-bit with alpha
        "u2c3": rng.integers(0, 65536, (4, 3, 3)).astype(np.uint16),
        "f8c2": rng.standard_normal((3, 5, 2)),                    # double, 2 channels
        "i2": rng.integers(-3000, 3000, (4, 4)).astype(np.int16),
        "u4": rng.integers(0, 1 << 31, (3, 4)).astype(np.uint32),
        "i8": rng.integers(-1 << 40, 1 << 40, (2, 3)).astype(np.int64),
    }


FLIPS = ["leftright", "topdown", "transpose", "r90", "r180", "r270", "posetrans",
         "xy", "Xy", "xY", "XY", "yx", "Yx", "yX", "YX"]
FLIPS3 = ["xyz", "yxz", "xzy", "zxy", "zyx", "yzx"]


def cases():
    """(reader, name template, IIO_TRANS) with {f} the input file"""
    c = []
    for k in ("f1", "f3", "u1c3", "f8c2"):
        c.append(("split", "{%s}" % k, ""))
        for fl in FLIPS:
            c.append(("split", "TRANS[flip=%s]:{%s}" % (fl, k), ""))
    for k in ("f3", "u1c3", "f8c2"):
        for fl in FLIPS3:
            c.append(("split", "TRANS[flip=%s]:{%s}" % (fl, k), ""))
    for k in ("f1", "f3", "i2", "u4", "i8", "u1c3", "f8c2"):
        for t in ("x=2", "y=1", "w=3", "h=2", "x=-2", "y=-3", "w=9", "h=6",
                  "x=1,y=1,w=3,h=2", "x=-1,y=-1"):
            c.append(("split", "TRANS[%s]:{%s}" % (t, k), ""))
    c += [
        ("split", "TRANS[flip=r90,x=1]:{f3}", ""),
        ("split", "TRANS[x=1]:TRANS[flip=r90]:{f1}", ""),
        ("split", "TRANS[flip=leftright,nonsense,=3,x=1]:{f1}", ""),
        ("split", "TRANS[flip=unknown]:{f1}", ""),
        ("split", "TRANS[pipe=cat]:{u1c3}", ""),
        ("split", "TRANS[pipe=cat,flip=topdown]:{f3}", ""),
        ("split", "{f1}", "flip=topdown"),
        ("split", "{f3}", "x=1,flip=r90"),
        ("split", "TRANS[flip=leftright]:{f1}", "x=1"),
        ("split", "TRANS[flip=leftright]:{f1}", "flip=transpose"),
        ("split", "zero:7x5", ""),
        ("split", "zero:4x3,2", ""),
        ("split", "one:6x2", ""),
        ("split", "one:3x4,3", ""),
        ("split", "constant:2.5:3x4", ""),
        ("split", "constant:-1e3:2x2", ""),
        ("split", "zero:7x5", "flip=transpose"),       # no IIO_TRANS for them
        ("split", "TRANS[flip=transpose]:one:3x2", ""),
        ("split", "TRANS[x=-1]:constant:4:2x3", ""),
        ("mask", "{f1}", ""),
        ("mask", "{u1c3}", ""),
        ("mask", "{u1c4}", ""),
        ("mask", "{u2c3}", ""),
        ("mask", "{f3}", ""),
        ("mask", "TRANS[flip=r90]:{u1c3}", ""),
        ("mask", "TRANS[x=-1,y=-1]:{u1c3}", ""),
        ("mask", "one:3x4,3", ""),
        ("mask", "one:5x2,4", ""),
        ("mask", "{f1}", "flip=leftright"),
    ]
    return c


def main():
    lib = ctypes.CDLL(REF_SO)
    for fn in ("iio_read_image_float_split", "iio_read_image_float"):
        getattr(lib, fn).restype = ctypes.POINTER(ctypes.c_float)
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    ins = inputs()
    out = {"in_" + k: v for k, v in ins.items()}
    table = []
    with tempfile.TemporaryDirectory() as d:
        paths = {}
        for k, v in ins.items():
            paths[k] = os.path.join(d, k + ".npy")
            np.save(paths[k], v)
        for i, (reader, tmpl, env) in enumerate(cases()):
            name = tmpl.format(**paths)
            os.environ.pop("IIO_TRANS", None)
            if env:
                os.environ["IIO_TRANS"] = env
            w, h, pd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(1)
            if reader == "split":
                p = lib.iio_read_image_float_split(name.encode(), ctypes.byref(w),
                                                   ctypes.byref(h), ctypes.byref(pd))
            else:
                p = lib.iio_read_image_float(name.encode(), ctypes.byref(w),
                                             ctypes.byref(h))
            a = np.ctypeslib.as_array(p, (pd.value, h.value, w.value)).copy()
            libc.free(ctypes.cast(p, ctypes.c_void_p))
            out["out_%d" % i] = a
            table.append({"reader": reader, "name": tmpl, "env": env})
        os.environ.pop("IIO_TRANS", None)
    out["cases"] = np.array(json.dumps(table))
    np.savez_compressed(os.path.join(HERE, "iio_names.npz"), **out)
    print("wrote", len(table), "cases")


if __name__ == "__main__":
    main()
