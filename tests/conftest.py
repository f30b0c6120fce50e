import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


class SmallCase:
    """Small configuration: 9x7x4 grid, 20k px over 5000-5500 A, 200 lines;
    15k observed pixels, vsini 10, RV 20, R 85000, SNR 100; 5 chunks of degree 4."""

    def __init__(self):
        from pasta_amd import synth
        from oracle import oracle as O

        self.spec = synth.small_grid_spec()
        self.labels = self.spec.labels()
        self.spectra, _ = synth.make_grid(self.spec)
        self.og = O.Grid(self.labels, self.spectra, None, self.spec.log10_first,
                         self.spec.log10_step)
        self.wl = synth.observed_wavelengths(self.spec, 15_000, 450)
        self.odisp = O.Dispersion(self.og, O.DISP_FIXED, self.wl, resolution=85_000.0)
        self.true = np.array([6123.0, -0.3, 4.21, 10.0, 20.0])
        model = O.produce_model(self.og, self.odisp, self.true)
        self.flux, self.var = synth.observe(model, self.wl, 100.0, 2000)
        self.ocont = O.Continuum("chunk", len(self.wl), wl=self.wl, n_chunks=5, p_order=4,
                                 overlap=0.2)


_small = None


@pytest.fixture(scope="session")
def small():
    global _small
    if _small is None:
        _small = SmallCase()
    return _small


@pytest.fixture(scope="session")
def gpu_small(small):
    import pasta_amd as pasta

    wlg = pasta.WlGrid(small.spec.log10_first, small.spec.log10_step, small.spec.n_pix, log=True)
    interp = pasta.InMemInterpolator.from_arrays(small.spectra, small.labels, wlg)
    return pasta, wlg, interp


_full = None


@pytest.fixture(scope="session")
def full():
    """The full-size grid (25 x 16 x 21 models of 200,000 pixels over 3800-9000 A, 3.36 GB,
    filled in HBM with the seeded 5,000-line recipe) and the oracle on the same rows read
    back from HBM."""
    global _full
    if _full is None:
        import pasta_amd as pasta
        from pasta_amd import synth
        from oracle import oracle as O

        spec = synth.full_grid_spec()
        wlg = pasta.WlGrid(spec.log10_first, spec.log10_step, spec.n_pix, log=True)
        interp = pasta.InMemInterpolator.empty(spec.labels(), wlg)
        synth.fill_grid_gpu(interp, spec)
        rows = synth.read_grid(interp)
        og = O.Grid(spec.labels(), rows, None, spec.log10_first, spec.log10_step)
        print(f"full grid {rows.shape} in HBM and on the host", flush=True)
        _full = (pasta, synth, spec, wlg, interp, og)
    return _full
