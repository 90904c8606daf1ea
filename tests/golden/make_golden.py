"""Generate the golden vectors in tests/golden/ with scipy.signal.

The reference's own tests (tests/torchscience/signal_processing/filter/
test__sosfilt.py and test__lfilter.py) compare against scipy.signal at run
time with unseeded inputs and commit no data.  This script fixes the same
cases with seeded inputs, runs scipy.signal.sosfilt / lfilter / sosfilt_zi /
lfilter_zi (scipy is the reference tests' oracle) and stores inputs and
outputs as .npz.  scipy's zi layout is (n_sections, batch..., 2); the
reference's is (batch..., n_sections, 2), so zi/zf are transposed here.

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy
import scipy.signal as ss

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from torchscience_amd.signal_processing.filter._design import (  # noqa: E402
    butterworth_sos, chebyshev1_sos)


def _rng(seed):
    return np.random.default_rng(seed)


def sos_cases():
    """(name, sos, x-shape, axis, with_zi) mirroring test__sosfilt.py."""
    return [
        ("butter5_025", ss.butter(5, 0.25, output="sos"), (200,), -1, False),        # :13
        ("cheby1_6", ss.cheby1(6, 0.5, 0.3, output="sos"), (150,), -1, False),       # :26
        ("ellip4", ss.ellip(4, 0.5, 40, 0.2, output="sos"), (100,), -1, False),      # :39
        ("butter2_single", ss.butter(2, 0.3, output="sos"), (100,), -1, False),      # :52
        ("butter5_high", ss.butter(5, 0.4, btype="high", output="sos"), (150,), -1, False),  # :65
        ("butter4_band", ss.butter(4, [0.2, 0.4], btype="band", output="sos"), (200,), -1, False),  # :78
        ("butter3_stop", ss.butter(3, [0.3, 0.5], btype="bandstop", output="sos"), (150,), -1, False),  # :91
        ("butter6_long", ss.butter(6, 0.2, output="sos"), (10000,), -1, False),      # :106
        ("butter12_many", ss.butter(12, 0.2, output="sos"), (200,), -1, False),     # :440
        ("butter3_single_sample", ss.butter(3, 0.3, output="sos"), (1,), -1, False),  # :427
        ("butter5_zi", ss.butter(5, 0.25, output="sos"), (100,), -1, True),          # :126
        ("butter4_axis_m1", ss.butter(4, 0.3, output="sos"), (10, 100), -1, False),   # :190
        ("butter3_axis0", ss.butter(3, 0.3, output="sos"), (100, 10), 0, False),      # :205
        ("butter2_axis1", ss.butter(2, 0.4, output="sos"), (5, 100, 8), 1, False),    # :221
        ("butter4_batch_zi", ss.butter(4, 0.3, output="sos"), (10, 100), -1, True),   # :240
        ("ragged_tail", ss.butter(4, 0.3, output="sos"), (3, 4097), -1, True),
        # reference-style design (even gain split, _zpk_to_sos.py:153-154): scipy's
        # own cheby1 sos puts the gain in section 0 and drives internal states
        # to 1e28, which no FP64 implementation reproduces beyond ~1e-7
        ("cheby1_16_near_unit", chebyshev1_sos(16, 0.01, 1.0), (2, 9000), -1, True),
    ]


def lf_cases():
    """(name, b, a, x-shape, with_zi) mirroring test__lfilter.py."""
    b5, a5 = ss.butter(5, 0.25)
    bc, ac = ss.cheby1(4, 0.5, 0.3)
    be, ae = ss.ellip(3, 0.5, 40, 0.2)
    bl, al = ss.butter(4, 0.2)
    b4z, a4z = ss.butter(4, 0.25)
    return [
        ("simple_iir", np.array([0.5, 0.5]), np.array([1.0, -0.3]), (100,), False),  # :12
        ("butter5", b5, a5, (200,), False),                                           # :26
        ("cheby1_4", bc, ac, (150,), False),                                          # :40
        ("ellip3", be, ae, (100,), False),                                            # :54
        ("fir", np.array([0.2, 0.3, 0.3, 0.2]), np.array([1.0]), (100,), False),     # :68
        ("long", bl, al, (10000,), False),                                            # :81
        ("butter4_zi", b4z, a4z, (100,), True),                                       # :98
        ("unnormalized", np.array([1.0, 1.0]), np.array([2.0, -0.6]), (100,), False),  # :335
        ("b_longer", np.array([0.1, 0.2, 0.3, 0.2, 0.1]), np.array([1.0, -0.5]), (100,), False),  # :348
        ("a_longer", np.array([0.5]), np.array([1.0, -0.5, 0.2, -0.1]), (100,), False),  # :361
        ("order2_batch_zi", np.array([0.2, 0.4, 0.2]), np.array([1.0, -0.6, 0.3]), (6, 300), True),
    ]


def ff_cases():
    """(name, kind, filter, x-shape, axis, padtype, padlen) mirroring
    test__sosfiltfilt.py and test__filtfilt.py (kind "sos": filter = sos;
    kind "ba": filter = (b, a))."""
    sb = ss.butter
    return [
        ("sos_butter5", "sos", sb(5, 0.25, output="sos"), (200,), -1, "odd", None),      # :13
        ("sos_cheby1_4", "sos", ss.cheby1(4, 0.5, 0.3, output="sos"), (150,), -1, "odd", None),  # :29
        ("sos_ellip3", "sos", ss.ellip(3, 0.5, 40, 0.2, output="sos"), (100,), -1, "odd", None),  # :45
        ("sos_bessel4", "sos", ss.bessel(4, 0.3, output="sos"), (100,), -1, "odd", None),  # :61
        ("sos_pad_even", "sos", sb(3, 0.3, output="sos"), (100,), -1, "even", None),     # :160
        ("sos_pad_const", "sos", sb(3, 0.3, output="sos"), (100,), -1, "constant", None),  # :176
        ("sos_pad_none", "sos", sb(3, 0.3, output="sos"), (100,), -1, None, None),       # :194
        ("sos_padlen20", "sos", sb(3, 0.3, output="sos"), (100,), -1, "odd", 20),        # :214
        ("sos_padlen0", "sos", sb(3, 0.3, output="sos"), (100,), -1, "odd", 0),          # :230
        ("sos_axis_m1", "sos", sb(3, 0.3, output="sos"), (10, 100), -1, "odd", None),    # :250
        ("sos_axis0", "sos", sb(3, 0.3, output="sos"), (100, 10), 0, "odd", None),       # :268
        ("sos_long", "sos", sb(4, 0.2, output="sos"), (10000,), -1, "odd", None),        # :354
        ("sos_ragged", "sos", sb(6, [0.05, 0.3], btype="band", output="sos"), (3, 9001), -1,
         "odd", None),
        ("ba_butter5", "ba", ss.butter(5, 0.25), (200,), -1, "odd", None),                # :13
        ("ba_cheby1_4", "ba", ss.cheby1(4, 0.5, 0.3), (150,), -1, "odd", None),          # :30
        ("ba_ellip3", "ba", ss.ellip(3, 0.5, 40, 0.2), (100,), -1, "odd", None),         # :47
        ("ba_fir", "ba", (np.array([0.2, 0.3, 0.3, 0.2]), np.array([1.0])), (100,), -1,
         "odd", None),                                                                    # :64
        ("ba_pad_even", "ba", ss.butter(3, 0.3), (100,), -1, "even", None),              # :169
        ("ba_pad_const", "ba", ss.butter(3, 0.3), (100,), -1, "constant", None),         # :186
        ("ba_pad_none", "ba", ss.butter(3, 0.3), (100,), -1, None, None),                # :205
        ("ba_padlen20", "ba", ss.butter(3, 0.3), (100,), -1, "odd", 20),                 # :226
        ("ba_padlen0", "ba", ss.butter(3, 0.3), (100,), -1, "odd", 0),                   # :243
        ("ba_axis_m1", "ba", ss.butter(3, 0.3), (10, 100), -1, "odd", None),             # :264
        ("ba_order2", "ba", ss.butter(2, 0.3), (4, 300), -1, "odd", None),
    ]


def main():
    out = {}
    manifest = {"scipy": scipy.__version__, "numpy": np.__version__, "cases": []}
    for i, (name, sos, shape, axis, with_zi) in enumerate(sos_cases()):
        x = _rng(100 + i).standard_normal(shape)
        rec = {"sos": sos, "x": x, "axis": np.array(axis)}
        if with_zi:
            batch = tuple(d for j, d in enumerate(shape) if j != axis % len(shape))
            zi = _rng(200 + i).standard_normal(batch + (sos.shape[0], 2))
            zi_sp = np.moveaxis(zi, -2, 0)  # (S, batch..., 2)
            y, zf_sp = ss.sosfilt(sos, x, axis=axis, zi=zi_sp)
            rec["zi"] = zi
            rec["zf"] = np.moveaxis(zf_sp, 0, -2)
        else:
            y = ss.sosfilt(sos, x, axis=axis)
        rec["y"] = y
        for k, v in rec.items():
            out[f"sos/{name}/{k}"] = np.asarray(v)
        manifest["cases"].append({"kind": "sosfilt", "name": name})
    for i, (name, b, a, shape, with_zi) in enumerate(lf_cases()):
        x = _rng(300 + i).standard_normal(shape)
        rec = {"b": b, "a": a, "x": x}
        if with_zi:
            order = max(len(a), len(b)) - 1
            zi = _rng(400 + i).standard_normal(shape[:-1] + (order,))
            y, zf = ss.lfilter(b, a, x, zi=zi)
            rec["zi"], rec["zf"] = zi, zf
        else:
            y = ss.lfilter(b, a, x)
        rec["y"] = y
        for k, v in rec.items():
            out[f"lf/{name}/{k}"] = np.asarray(v)
        manifest["cases"].append({"kind": "lfilter", "name": name})
    for i, (name, kind, filt, shape, axis, padtype, padlen) in enumerate(ff_cases()):
        x = _rng(500 + i).standard_normal(shape)
        rec = {"x": x, "axis": np.array(axis),
               "padtype": np.array("none" if padtype is None else padtype),
               "padlen": np.array(-1 if padlen is None else padlen)}
        if kind == "sos":
            rec["sos"] = filt
            y = ss.sosfiltfilt(filt, x, axis=axis, padtype=padtype, padlen=padlen)
        else:
            rec["b"], rec["a"] = filt
            y = ss.filtfilt(filt[0], filt[1], x, axis=axis, padtype=padtype, padlen=padlen)
        rec["y"] = y
        for k, v in rec.items():
            out[f"ff/{name}/{k}"] = np.asarray(v)
        manifest["cases"].append({"kind": "sosfiltfilt" if kind == "sos" else "filtfilt",
                                  "name": name})
    # steady-state helpers
    for name, sos in [("butter5", ss.butter(5, 0.25, output="sos")),
                      ("butter6", ss.butter(6, 0.2, output="sos"))]:
        out[f"zi/sos/{name}/sos"] = sos
        out[f"zi/sos/{name}/zi"] = ss.sosfilt_zi(sos)
    b, a = ss.butter(5, 0.2)
    out["zi/lf/butter5/b"], out["zi/lf/butter5/a"] = b, a
    out["zi/lf/butter5/zi"] = ss.lfilter_zi(b, a)
    # configuration 1: order-8 Butterworth lowpass at 0.125 (FP64 design with
    # the reference's pairing), x (16, 4096) from torch.Generator seed 0.
    import torch
    g = torch.Generator().manual_seed(0)
    x1 = torch.randn(16, 4096, dtype=torch.float64, generator=g).numpy()
    sos1 = butterworth_sos(8, 0.125)
    out["c1/sos"] = sos1
    out["c1/y"] = ss.sosfilt(sos1, x1, axis=-1)
    out["c1/x_checksum"] = np.array([x1.sum(), (x1 * x1).sum(), x1[3, 77], x1[15, 4095]])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
