"""Host bookkeeping of batch.calc_films_many (no GPU): the device entry point
emu_film_modes_batch is replaced by a numpy stand-in with its documented
contract (k_z of order order[b, i] at position i, both polarisation blocks;
Z = k / (n k_z) and k_z / (n k); Z = 0 where k_z is exactly 0), so that what
is checked is the Python side: the orders and wavevectors handed over, the
free-space reference's order, the specular order and propagating count, and
the one retry with k_parallel scaled by 1 - 1e-9 for a light whose orders
include a grazing one (Anallo.calc_kz's rule, reference mode_calcs.py:
147-162)."""
import numpy as np
import pytest

PERIOD = 600.0


class _FakeContext:
    def __init__(self, graze_first_call=None):
        self.calls = []
        self.graze = graze_first_call

    def film_modes(self, n, wl_norm, alpha, beta, order):
        n, wl_norm = np.asarray(n), np.asarray(wl_norm)
        k = np.complex128(2 * np.pi) * n / wl_norm
        kz_u = np.sqrt(k[:, None] ** 2 - alpha ** 2 - beta ** 2)
        if self.graze is not None and not self.calls:
            kz_u[self.graze, 0] = 0.0  # a grazing order on the first evaluation
        self.calls.append((alpha.copy(), beta.copy()))
        kz = np.take_along_axis(kz_u, order, axis=1)
        with np.errstate(divide='ignore', invalid='ignore'):
            zte = np.where(kz == 0, 0, (1 / n)[:, None] * k[:, None] / kz)
            ztm = np.where(kz == 0, 0, (1 / n)[:, None] * kz / k[:, None])
        return np.concatenate((kz, kz), axis=1), np.concatenate((zte, ztm), axis=1)


def _setup(monkeypatch, fake):
    from emustack_amd import EMUstack
    monkeypatch.setattr(EMUstack, 'context', lambda: fake)


def _lights(theta=12.0):
    from emustack_amd import objects
    return [objects.Light(wl, max_order_PWs=3, theta=theta, phi=25.0) for wl in (450.0, 610.0, 880.0)]


@pytest.mark.parametrize('world_1d', [False, True])
def test_films_many_equals_calc_modes(monkeypatch, world_1d):
    from emustack_amd import objects, materials
    from emustack_amd.batch import calc_films_many
    film = objects.ThinFilm(PERIOD, height_nm=80.0, world_1d=world_1d,
                            material=materials.Material(2.1 + 0.2j), loss=True)
    fake = _FakeContext()
    _setup(monkeypatch, fake)
    got = calc_films_many(film, _lights())
    assert len(fake.calls) == 1
    for f, light in zip(got, _lights()):
        ref = film.calc_modes(light)
        np.testing.assert_allclose(f.k_z, ref.k_z, rtol=1e-14, atol=0)
        np.testing.assert_allclose(f.Z(), ref.Z(), rtol=1e-14, atol=0)
        np.testing.assert_array_equal(f.sort_order, ref.sort_order)
        assert f.specular_order == ref.specular_order
        assert f.num_prop_pw_per_pol == ref.num_prop_pw_per_pol
        np.testing.assert_array_equal(f.alphas, ref.alphas)
        np.testing.assert_array_equal(np.broadcast_to(f.betas, f.alphas.shape),
                                      np.broadcast_to(ref.betas, ref.alphas.shape))


def test_grazing_light_is_evaluated_again_with_scaled_k_parallel(monkeypatch, capsys):
    from emustack_amd import objects, materials
    from emustack_amd.batch import calc_films_many
    film = objects.ThinFilm(PERIOD, height_nm=80.0, material=materials.Material(1.5 + 0j), loss=False)
    fake = _FakeContext(graze_first_call=1)
    _setup(monkeypatch, fake)
    lights = _lights()
    got = calc_films_many(film, lights)
    assert len(fake.calls) == 2
    a0, a1 = fake.calls[0][0], fake.calls[1][0]
    np.testing.assert_array_equal(a1[[0, 2]], a0[[0, 2]])  # the other lights unchanged
    kp = lights[1].k_pll * PERIOD
    px = np.round((a0[1] - kp[0]) / (2 * np.pi)).astype(int)
    np.testing.assert_allclose(a1[1], (1 - 1e-9) * kp[0] + 2 * np.pi * px, rtol=0, atol=1e-12)
    assert 'Wood anomaly' in capsys.readouterr().out
    assert np.all(got[1].k_z != 0)
