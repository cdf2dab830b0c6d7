"""Write ``shade_cases.json``: the inputs and expected images of the reference's own
``tf.shade`` tests, as data.

Every case restates one assertion of datashader 0.19.1's
``datashader/tests/test_transfer_functions.py`` (line numbers below): the plane the test
builds, the arguments of the ``tf.shade`` call and the uint32 image it asserts.  The
``check_span`` sequences (123-170) are unrolled into one case per call.  Cases the
reference compares with ``assert_image_close(..., tolerance=1)`` carry ``"tol": 1``.
Callable ``how`` / matplotlib ``cmap`` cases are left out (this backend does not take
them).  The generated JSON is committed; run this again only to add a case.

Usage: python tests/golden/make_shade_golden.py
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "shade_cases.json")
T = "datashader/tests/test_transfer_functions.py"


def enc(a):
    a = np.asarray(a)
    if a.dtype.kind == "f":
        return [None if math.isnan(v) else float(v) for v in a.ravel()]
    if a.dtype.kind == "b":
        return [int(v) for v in a.ravel()]
    return [int(v) for v in a.ravel()]


def plane(a):
    a = np.asarray(a)
    return {"dtype": a.dtype.str[1:], "shape": list(a.shape), "values": enc(a)}


def build_agg():
    """build_agg (23-38)."""
    a = np.arange(10, 19, dtype="u4").reshape((3, 3))
    a[[0, 1, 2], [0, 1, 2]] = 0
    b = np.arange(10, 19, dtype="f4").reshape((3, 3))
    b[[0, 1, 2], [0, 1, 2]] = np.nan
    c = np.arange(10, 19, dtype="f8").reshape((3, 3))
    c[[0, 1, 2], [0, 1, 2]] = np.nan
    d = np.arange(10, 19, dtype="u4").reshape((3, 3))
    d[[0, 1, 2, 2], [0, 1, 2, 1]] = 1
    return dict(a=a, b=b, c=c, d=d)


SOL = {  # solution_lists (89-103)
    "log": [[0, 4291543295, 4286741503], [4283978751, 0, 4280492543],
            [4279242751, 4278190335, 0]],
    "cbrt": [[0, 4291543295, 4284176127], [4282268415, 0, 4279834879],
             [4278914047, 4278190335, 0]],
    "linear": [[0, 4291543295, 4289306879], [4287070463, 0, 4282597631],
               [4280361215, 4278190335, 0]],
}
EQ_SOL = [[0, 4291543295, 4288846335], [4286149631, 0, 4283518207],
          [4280821503, 4278190335, 0]]  # 109-111
EQ_RESCALE = {  # 113-120
    "a": [[0, 4289306879, 4287070463], [4284834047, 0, 4282597631],
          [4280361215, 4278190335, 0]],
    "b": [[0, 4289306879, 4287070207], [4284834047, 0, 4282597375],
          [4280361215, 4278190335, 0]],
}
EQ_RESCALE["c"] = EQ_RESCALE["b"]

cases = []


def col(c):
    """A colour as JSON: a name / hex string, or {"tuple": [r, g, b]}."""
    return c if isinstance(c, str) else {"tuple": [int(v) for v in c]}


def add(name, cite, data, expected, tol=0, **kw):
    if "cmap" in kw:
        cm = kw["cmap"]
        kw["cmap"] = [col(c) for c in cm] if isinstance(cm, list) else col(cm)
    if "color_key" in kw:
        ck = kw["color_key"]
        kw["color_key"] = ({"dict": [[k, col(v)] for k, v in ck.items()]}
                           if isinstance(ck, dict) else [col(c) for c in ck])
    cases.append({"name": name, "cite": f"{T}:{cite}", "data": plane(data),
                  "expected": np.asarray(expected, dtype=np.uint32).ravel().tolist(),
                  "tol": tol, "kwargs": kw})


def check_span(tag, cite, x, cmap, how, sol):
    """check_span (123-170) unrolled."""
    x = x.copy()
    sol = np.array(sol, dtype=np.uint32)
    add(f"{tag}/span-none", cite, x, sol, cmap=cmap, how=how, span=None)
    add(f"{tag}/span-float", cite, x, sol, cmap=cmap, how=how, span=[11.0, 17.0])
    x[0, 1] = 10
    add(f"{tag}/below-span", cite, x, sol, cmap=cmap, how=how, span=[11.0, 17.0])
    x[2, 1] = 18
    add(f"{tag}/above-span", cite, x, sol, cmap=cmap, how=how, span=[11.0, 17.0])
    x[0, 1] = 0 if x.dtype.kind in ("i", "u") else np.nan
    sol = sol.copy()
    sol[0, 1] = sol[0, 0]
    add(f"{tag}/masked-low", cite, x, sol, cmap=cmap, how=how, span=[11.0, 17.0])
    x[2, 1] = 0 if x.dtype.kind in ("i", "u") else np.nan
    sol[2, 1] = sol[0, 0]
    add(f"{tag}/masked-high", cite, x, sol, cmap=cmap, how=how, span=[11.0, 17.0])


agg = build_agg()
pink_red = ["pink", "red"]
for attr in "abc":
    x = agg[attr]
    for span in (None, [11, 17], [11.0, 17.0]):
        for how in ("log", "cbrt", "linear"):
            add(f"test_shade/{attr}/{how}/span={span}", "173-211", x, SOL[how],
                cmap=pink_red, how=how, span=span)
    add(f"test_shade/{attr}/eq_hist", "175-211", x, EQ_SOL, cmap=pink_red, how="eq_hist")
    add(f"test_shade/{attr}/eq_hist-rescale", "175-211", x, EQ_RESCALE[attr], tol=1,
        cmap=pink_red, how="eq_hist", rescale_discrete_levels=True)
    for how in ("linear", "log", "cbrt"):
        for cm in (["pink", "red"], ["#FFC0CB", "#FF0000"]):
            check_span(f"test_span_cmap_list/{attr}/{how}/{cm[0]}", "213-224", x, cm, how,
                       SOL[how])

single = [[0, 671088640, 1946157056], [2701131776, 0, 3640655872],
          [3976200192, 4278190080, 0]]  # 233-235
for cm in ("black", (0, 0, 0), "#000000"):
    check_span(f"test_span_cmap_single/{cm}", "227-239", agg["a"], cm, "log", single)
    add(f"test_shade_cmap_non_categorical_alpha/{cm}", "285-292", agg["a"], single, cmap=cm,
        how="log")

boolean = ~np.eye(3, dtype="bool")  # 260-272
for how in ("log", "cbrt", "linear", "eq_hist"):
    add(f"test_shade_bool/{how}", "260-272", boolean,
        np.where(boolean, 4278190335, 0).astype("uint32"), cmap=pink_red, how=how)

add("test_shade_cmap", "275-282", agg["a"],
    [[0, 4278190335, 4278236489], [4280344064, 0, 4289091584],
     [4292225024, 4294901760, 0]], cmap=["red", (0, 255, 0), "#0000FF"], how="log")

for rescale, sol in ((False, [[0xffcbc0ff, 0xffa299ff, 0xff7973ff],
                              [0xff514cff, 0xffcbc0ff, 0xff2826ff],
                              [0xff0000ff, 0xffcbc0ff, 0xffcbc0ff]]),
                     (True, [[0xff8981ff, 0xff6d67ff, 0xff524dff],
                             [0xff3633ff, 0xff8981ff, 0xff1b19ff],
                             [0xff0000ff, 0xff8981ff, 0xff8981ff]])):
    add(f"test_shade_rescale_discrete_levels/{rescale}", "536-551", agg["d"], sol,
        cmap=pink_red, how="eq_hist", rescale_discrete_levels=rescale)

# --- categorical (313-496) -----------------------------------------------------------------
colors = [(255, 0, 0), "#0000FF", "orange"]
cat = np.array([[(0, 12, 0), (3, 0, 3)], [(12, 12, 12), (24, 0, 0)]], dtype="u4")
C = "313-496"
for how, sol in (("log", [[2583625728, 335565567], [4283774890, 3707764991]]),
                 ("cbrt", [[2650734592, 335565567], [4283774890, 3657433343]]),
                 ("linear", [[1140785152, 335565567], [4283774890, 2701132031]])):
    add(f"test_shade_category/{how}", C, cat, sol, color_key=colors, how=how, min_alpha=20)
for span, sol in (([50, 100], [[16711680, 21247], [5584810, 255]]),
                  ([0, 2], [[4294901760, 4278211327], [4283774890, 4278190335]]),
                  ([6, 36], [[872349696, 21247], [4283774890, 2566914303]]),
                  ([0, 72], [[721354752, 352342783], [2136291242, 1426063615]])):
    add(f"test_shade_category/span={span}", C, cat, sol, color_key=colors, how="linear",
        min_alpha=0, span=span)
cat0 = np.array([[(0, 0, 0), (3, 0, 3)], [(12, 12, 12), (24, 0, 0)]], dtype="u4")
add("test_shade_category/empty-pixel", C, cat0, [[5584810, 335565567], [4283774890, 2701132031]],
    color_key=colors, how="linear", min_alpha=20)
add("test_shade_category/empty-pixel-span", C, cat0,
    [[5584810, 335565567], [4283774890, 2701132031]], color_key=colors, how="linear",
    min_alpha=20, span=[6, 36])
neg = np.array([[(0, -30, 0), (18, 0, -18)], [(-2, 2, -2), (-18, 9, 12)]], dtype="i4")
add("test_shade_category/signed", C, neg, [[335565567, 3914667690], [3680253090, 4285155988]],
    color_key=colors, how="linear", min_alpha=20)
add("test_shade_category/signed-span", C, neg, [[335565567, 341120682], [341587106, 4285155988]],
    color_key=colors, how="linear", min_alpha=20, span=[0, 3])
add("test_shade_category/signed-baseline", C, neg,
    [[341129130, 3909091583], [3679795114, 4278232575]], color_key=colors, how="linear",
    min_alpha=20, color_baseline=9)
allneg = np.array([[(0, -30, 0), (-18, 0, -18)], [(-2, -2, -2), (-18, 0, 0)]], dtype="i4")
add("test_shade_category/negative", C, allneg, [[1124094719, 344794225], [4283774890, 2708096148]],
    color_key=colors, how="linear", min_alpha=20)
add("test_shade_category/negative-span", C, allneg,
    [[335565567, 344794225], [341129130, 342508692]], color_key=colors, how="linear",
    min_alpha=20, span=[6, 36])
add("test_shade_zeros", "520-533", np.zeros((2, 2, 3), dtype="u4"),
    [[5584810, 5584810], [5584810, 5584810]], color_key=colors, how="linear", min_alpha=0)
catd = np.array([[[1, 2], [0, 1]], [[0, 0], [0, 0]], [[1, 0], [3, 0]], [[1, 0], [2, 1]]],
                dtype="u4")
add("test_shade_rescale_discrete_levels_categorical", "554-566", catd,
    [[0xff845c70, 0x6fb87e37], [0x006a4c8d, 0x006a4c8d], [0x6f1c1ae4, 0xff1c1ae4],
     [0x6f1c1ae4, 0xff503baa]], how="eq_hist", rescale_discrete_levels=True)

# --- 2D colour key (1181-1197) --------------------------------------------------------------
key_data = np.array([[0, 0, 0, 0, 0], [0, 1, 1, 1, 0], [0, 2, 2, 2, 0], [0, 3, 3, 3, 0],
                     [0, 0, 0, 0, 0]], dtype="uint32")
add("test_shade_with_discrete_color_key", "1181-1197", key_data,
    [[0, 0, 0, 0, 0], [0, 4294967295, 4294967295, 4294967295, 0],
     [0, 4286578816, 4286578816, 4286578816, 0], [0, 4278255615, 4278255615, 4278255615, 0],
     [0, 0, 0, 0, 0]], color_key={1: "white", 2: "purple", 3: "yellow"})

with open(OUT, "w") as f:
    json.dump({"source": "datashader 0.19.1 " + T, "cases": cases}, f, indent=None)
    f.write("\n")
print(f"{len(cases)} cases -> {OUT}")
