"""Pins the oracle's broadening kernels, overlap-add convolution and RV resample
(reference src/convolve_rv.rs) against closed forms: the textbook limb-darkened
rotation profile (Gray, The Observation and Analysis of Stellar Photospheres),
direct FP64 convolution (numpy.convolve), the normalised Gaussian, and
closed-form linear interpolation."""
import numpy as np
import pytest

from oracle import oracle as O

LN10 = np.log(10.0)


@pytest.mark.parametrize("vsini", [5.0, 10.0, 57.3, 150.0, 300.0])
@pytest.mark.parametrize("step", [2.0696e-6, 1.8723e-6])
def test_rotation_kernel_closed_form(vsini, step):
    dvelo = LN10 * step
    k = O.rotation_kernel(vsini, dvelo)
    n = len(k)
    assert n % 2 == 1
    vrot = vsini / 299792.0
    # the profile is supported on |v| < vrot; the reference sizes the kernel as
    # ceil(2 vrot/dv * sqrt(1 + (4(eps-1)/(pi eps))^2)) + 2 taps, made odd (convolve_rv.rs:443-449)
    half = (n - 1) / 2 * dvelo
    widen = np.sqrt(1 + (4 * (0.6 - 1) / (np.pi * 0.6)) ** 2)
    assert half >= vrot
    assert abs(n - (2 * vrot / dvelo * widen + 2)) <= 2
    # Gray's rotation profile G(v) with linear limb darkening eps = 0.6, normalised
    eps = 0.6
    v = (np.arange(n) - (n - 1) / 2) * dvelo
    x2 = 1 - (v / vrot) ** 2
    g = np.where(x2 > 0, 2 * (1 - eps) * np.sqrt(np.clip(x2, 0, None)) + np.pi * eps / 2 * x2, 0.0)
    g /= g.sum()
    np.testing.assert_allclose(k, g, rtol=2e-6, atol=1e-9)
    assert abs(float(k.astype(np.float64).sum()) - 1.0) < 1e-5
    np.testing.assert_allclose(k, k[::-1], rtol=1e-6, atol=1e-10)


@pytest.mark.parametrize("m", [3, 7, 21, 205, 611, 2047])
@pytest.mark.parametrize("l", [100, 2000, 20_000])
def test_oa_convolve_equals_direct(m, l):
    rng = np.random.default_rng(m * 7 + l)
    x = rng.uniform(0.2, 1.0, l).astype(np.float32)
    h = rng.uniform(0, 1, m).astype(np.float32)
    h /= h.sum()
    y = O.oa_convolve(x, h)
    ref = np.convolve(x.astype(np.float64), h.astype(np.float64))
    assert len(y) >= len(ref)
    # f32 FFT rounding; a 2047-tap kernel leaves a 2-sample block step, so each output
    # sums ~1000 block contributions
    tol = 1e-6 if m <= 611 else 1e-5
    np.testing.assert_allclose(y[: len(ref)], ref, rtol=tol, atol=tol * np.abs(ref).max())
    assert np.all(np.abs(y[len(ref):]) < 1e-6)


@pytest.mark.parametrize("R", [5_000.0, 60_000.0, 85_000.0])
def test_fixed_gaussian_kernel(R):
    step = 2.0696e-6
    k = O.fixed_gauss_kernel(R, step)
    sigma = 1.0 / (LN10 * 2.355 * R * step)
    n = len(k)
    assert n % 2 == 1 and n >= 6 * sigma and n <= 6 * sigma + 2
    x = np.arange(n) - (n - 1) / 2
    g = np.exp(-0.5 * (x / sigma) ** 2)
    np.testing.assert_allclose(k, g / g.sum(), rtol=1e-6)


def test_resample_weight_quantisation(small):
    """produce_model's resample (src/convolve_rv.rs:569-573) uses w = f32(idx) - f32(j):
    against the FP64 mode (exact weight) the difference stays within that quantisation."""
    og, od = small.og, small.odisp
    for lab in ([6123.0, -0.3, 4.21, 10.0, 20.0], [5500.0, 0.1, 4.6, 0.0, -77.0]):
        m32 = O.produce_model(og, od, lab)
        m64 = O.produce_model(og, od, lab, f64=True)
        idx = od_fidx = (np.log10(small.wl) - small.spec.log10_first) / small.spec.log10_step
        idx = od_fidx + np.log10(1 - lab[4] / 299792.0) / small.spec.log10_step
        j = np.floor(idx)
        dw = np.abs(idx.astype(np.float32).astype(np.float64) - j - (idx - j))
        # |m32 - m64| <= dw * |c[j+1]-c[j]| + f32 rounding; |c[j+1]-c[j]| <= 1 for normalised flux
        assert np.all(np.abs(m32 - m64) <= dw + 5e-6)


def test_produce_model_f32_close_to_f64_away_from_quantisation(small):
    og, od = small.og, small.odisp
    lab = [6000.0, -0.2, 4.1, 40.0, 0.0]  # rv = 0: idx = fidx, weights identical
    m32 = O.produce_model(og, od, lab)
    m64 = O.produce_model(og, od, lab, f64=True)
    assert np.max(np.abs(m32 - m64)) < 2e-3
    assert np.median(np.abs(m32 - m64)) < 1e-5


def test_range_errors(small):
    og, od = small.og, small.odisp
    with pytest.raises(O.OracleError) as e:
        O.produce_model(og, od, [6000.0, 0.0, 4.0, 10.0, 30000.0])
    assert e.value.code == -5
    with pytest.raises(O.OracleError) as e:
        O.produce_model(og, od, [6000.0, 0.0, 4.0, -1.0, 0.0])
    assert e.value.code == -3
