"""The resolve's unpackUnorm4x8 arithmetic (vkhr_amd/csrc/ppll.hip unorm8):
b / 255 for b = 0..255 as q = fl(b * r), e = fma(-q, 255, b), fl(fma(e, r, q))
with r = fl(1/255), against the correctly rounded quotient, in exact rational
arithmetic.  share/shaders/transparency/resolve.comp unpacks the stored
colour with unpackUnorm4x8, which divides by 255."""
from fractions import Fraction

import numpy as np


def round_f32(x: Fraction) -> Fraction:
    """x rounded to the nearest float32 (ties to even), exactly."""
    v = np.float32(float(x))  # within one float32 step of the answer
    cands = [np.nextafter(v, np.float32(-np.inf)), v, np.nextafter(v, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(np.float32(c).view(np.uint32)) & 1))
    return Fraction(float(best))


R = round_f32(Fraction(1, 255))


def division_step(b: int) -> Fraction:
    q = round_f32(b * R)
    e = round_f32(-q * 255 + b)  # fma: one rounding of the exact value
    return round_f32(e * R + q)


def test_round_f32_helper():
    assert round_f32(Fraction(1, 3)) == Fraction(float(np.float32(1.0) / np.float32(3.0)))
    # a tie between two float32 neighbours of 1 goes to the even one (1)
    assert round_f32(1 + Fraction(1, 2 ** 24)) == 1
    assert round_f32(1 + Fraction(3, 2 ** 24)) == 1 + Fraction(2, 2 ** 23)


def test_unorm8_division_step_is_the_correctly_rounded_quotient():
    for b in range(256):
        assert division_step(b) == round_f32(Fraction(b, 255)), b


def test_plain_reciprocal_product_is_not():
    # without the residual step the product by fl(1/255) misrounds some bytes,
    # so the test above distinguishes the two
    wrong = [b for b in range(256) if round_f32(b * R) != round_f32(Fraction(b, 255))]
    assert wrong
