"""The closed form of the percentage pre-sampler's golden-ratio sequence that
the device origin sampler uses (necsim-rust_amd/csrc/nh_origins.hip,
golden_q): the reference's f64 loop q_0 = 0.5, q_{j+1} = (q_j + INV_PHI) -
floor(q_j + INV_PHI) (origin_sampler/pre_sampler.rs:86-90) never rounds,
because 2^53 q_0 and 2^53 INV_PHI are both multiples of 16, so
q_j = ((2^52 + j P) mod 2^53) 2^-53 with P = INV_PHI 2^53.  Checked against
the f64 loop itself (IEEE float64, round to nearest even) over 2 million
steps, with the premise (P a multiple of 16) checked on its own: a
restatement with another constant or start would fail both.  CPU.
"""
import numpy as np

PHI = 6.180339887498949e-1  # pre_sampler.rs INV_PHI
Q_ONE = 1 << 53


def test_golden_sequence_closed_form_equals_the_f64_loop():
    P = int(np.ldexp(np.float64(PHI), 53))
    assert float(P) * 2.0**-53 == PHI
    assert P == 0x13C6EF372FE950 and P % 16 == 0 and (Q_ONE >> 1) % 16 == 0
    steps = 2_000_000
    # the reference's loop, in float64
    q = np.empty(steps, np.float64)
    x = 0.5
    for j in range(steps):
        x += PHI
        x -= float(np.floor(x))
        q[j] = x
    # the closed form, vectorised
    j = np.arange(1, steps + 1, dtype=np.uint64)
    Q = (np.uint64(Q_ONE >> 1) + j * np.uint64(P)) & np.uint64(Q_ONE - 1)
    assert np.array_equal(q, Q.astype(np.float64) * 2.0**-53)
    # the loop wraps about 62 % of the steps: the closed form is not trivially
    # the same as an unwrapped sum
    assert 0.6 < np.mean(np.diff(q) < 0) < 0.65
