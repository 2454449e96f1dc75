"""On-disk formats (CPU): FITS layout against the FITS standard, the
reference's data and checkpoint field names, the pixel-size and uv-scaling
rules of the imaging driver."""
import struct

import numpy as np
import pytest

from fhs import io as fio


def test_fits_layout_follows_the_standard(tmp_path):
    """Header of 80-character cards in 2880-byte records, big-endian IEEE
    data starting at the next record, NAXIS1 the fastest axis (MATLAB's
    first index), data padded with zeros to a whole record."""
    x = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4, order='F') * 0.5 - 3.0
    p = tmp_path / 'cube.fits'
    fio.fitswrite(x, p)
    raw = p.read_bytes()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode('ascii') for i in range(0, 2880, 80)]
    assert cards[0] == 'SIMPLE  =                    T / conforms to FITS standard'.ljust(80)
    assert cards[1] == 'BITPIX  =                  -64'.ljust(80)
    assert cards[2] == 'NAXIS   =                    3'.ljust(80)
    assert cards[3] == 'NAXIS1  =                    2'.ljust(80)
    assert cards[4] == 'NAXIS2  =                    3'.ljust(80)
    assert cards[5] == 'NAXIS3  =                    4'.ljust(80)
    assert cards[6] == 'END'.ljust(80)
    assert all(c == ' ' * 80 for c in cards[7:])
    body = raw[2880:]
    vals = struct.unpack('>24d', body[:24 * 8])
    assert vals[0] == x[0, 0, 0] and vals[1] == x[1, 0, 0] and vals[2] == x[0, 1, 0] and vals[6] == x[0, 0, 1]
    assert np.array_equal(np.array(vals), x.ravel(order='F'))
    assert body[24 * 8:] == b'\0' * (2880 - 24 * 8)


@pytest.mark.parametrize('dtype,bitpix', [(np.float64, -64), (np.float32, -32)])
def test_fits_round_trip(tmp_path, dtype, bitpix):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(17, 13, 3)).astype(dtype)
    p = tmp_path / 'x.fits'
    fio.fitswrite(x, p, dtype=dtype)
    assert f'BITPIX  = {bitpix:>20}' in p.read_bytes()[:2880].decode('ascii')
    y = fio.fitsread(p)
    assert y.dtype == dtype and y.shape == x.shape and np.array_equal(x, y)
    img = x[:, :, 0]
    fio.fitswrite(img, p, dtype=dtype)
    assert np.array_equal(fio.fitsread(p), img)


def test_fits_reader_applies_scaling_and_integer_types(tmp_path):
    """A 16-bit integer image with BZERO/BSCALE, built by hand to the standard."""
    vals = np.array([[1, -2], [300, 4]], dtype=np.int16)           # (NAXIS1, NAXIS2)
    cards = ['SIMPLE  =                    T', 'BITPIX  =                   16', 'NAXIS   =                    2',
             'NAXIS1  =                    2', 'NAXIS2  =                    2', 'BSCALE  =                  0.5',
             'BZERO   =                 10.0', 'END']
    head = ''.join(c.ljust(80) for c in cards).encode('ascii')
    head += b' ' * (-len(head) % 2880)
    body = vals.astype('>i2').tobytes(order='F')
    body += b'\0' * (-len(body) % 2880)
    p = tmp_path / 'int.fits'
    p.write_bytes(head + body)
    assert np.array_equal(fio.fitsread(p), vals * 0.5 + 10.0)


def test_data_file_fields(tmp_path):
    """data_ch_<k>.mat with the generator's field names
    (imaging/main_generate_data.m:345-364) written and read back."""
    rng = np.random.default_rng(1)
    M = 50
    f = dict(y0=rng.normal(size=M) + 1j * rng.normal(size=M), y=rng.normal(size=M) + 1j * rng.normal(size=M),
             l2bounds=np.array([3.0, 4.0]), sigma_noise=0.25, u=rng.normal(size=M), v=rng.normal(size=M),
             w=np.ones(M), frequency=1.4e9, maxProjBaseline=1234.5, nW=np.full(M, 0.25))
    p = tmp_path / 'data_ch_1.mat'
    fio.save_data_ch(p, **f)
    import scipy.io as sio
    raw = sio.loadmat(p)
    assert raw['y'].shape == (M, 1) and raw['u'].shape == (M, 1)            # MATLAB column vectors
    g = fio.load_data_ch(p)
    for k, v in f.items():
        assert np.allclose(g[k], v), k
    assert isinstance(g['frequency'], float) and 'nWimag' not in g
    with pytest.raises(ValueError):
        fio.save_data_ch(p, uvw=np.ones(3))


def test_mat_v73_is_refused(tmp_path):
    p = tmp_path / 'data_ch_2.mat'
    p.write_bytes(b'MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: x HDF5 schema 1.00 .'.ljust(512, b' '))
    with pytest.raises(ValueError, match='v7.3'):
        fio.load_data_ch(p)


def test_pixel_size_and_uv_scaling():
    """imaging/imaging.m:373-386 and util_load_data_gen_measurement_operator.m:89-91:
    at the nominal resolution (superresolution 1) the longest projected
    baseline maps to pi / 2 ... pi: u = maxProjBaseline -> pi / superresolution."""
    B = 2000.0
    for sr in (1.0, 1.5, 2.0):
        hb = fio.half_spatial_bandwidth(max_proj_baseline=B, superresolution=sr)
        assert np.isclose(hb, sr * B)
        u, v = fio.uv_to_radians(np.array([B, -B / 2]), np.array([B, 0.25 * B]), hb)
        assert np.allclose(u, [np.pi / sr, -np.pi / (2 * sr)])
        assert np.allclose(v, [-np.pi / sr, -np.pi / (4 * sr)])          # v changes sign
    ps = 2.5                                                             # arcsec
    assert np.isclose(fio.half_spatial_bandwidth(pixel_size=ps), (180 / np.pi) * 3600 / ps / 2)
    with pytest.raises(ValueError):
        fio.half_spatial_bandwidth(max_proj_baseline=B)


def test_checkpoint_fields(tmp_path):
    """<name>_rw<flag>.mat with the solver's field names
    (src/fhs/facetHyperSARA.m:1165-1205), cells of per-worker entries."""
    rng = np.random.default_rng(2)
    Ny, Nx, L = 8, 6, 3
    xsol, res, g = rng.random((Ny, Nx, L)), rng.normal(size=(Ny, Nx, L)), rng.normal(size=(Ny, Nx, L))
    eps = [np.array([1.0, 2.0]), np.array([3.0])]
    v2 = [rng.normal(size=5) + 1j * rng.normal(size=5), rng.normal(size=4) + 0j]
    proj = [np.zeros(5, complex), np.ones(4, complex)]
    tb = [np.array([0, 0]), np.array([7])]
    nr = [np.array([0.5, 0.25]), np.array([1.5])]
    v0, v1 = [rng.normal(size=(Ny * Nx, L))], [rng.normal(size=(40, L))]
    w0, w1 = [np.ones(L)], [np.ones(40)]
    p = tmp_path / 'ckpt_rw2.mat'
    fio.save_checkpoint(p, xsol, res, g, eps, v2, proj, tb, nr, v0, v1, w0, w1,
                        param={'pdfb_max_iter': 2000, 'gam': 1.0}, end_iter=np.arange(3.0), rel_val=[1e-3, 1e-4])
    c = fio.load_checkpoint(p)
    assert set(c) == {'param', 'res', 'g', 'xsol', 'epsilon', 'v2', 'proj', 't_block', 'norm_res', 'v0', 'v1',
                      'weights0', 'weights1', 'end_iter', 'rel_val'}
    assert np.array_equal(c['xsol'], xsol) and c['xsol'].shape == (Ny, Nx, L)
    assert np.allclose(c['v2'][0], v2[0]) and np.allclose(c['v2'][1], v2[1])
    assert np.allclose(c['epsilon'][0], eps[0]) and np.allclose(c['v0'][0], v0[0])
    assert int(c['param']['pdfb_max_iter'][0, 0].squeeze()) == 2000


def test_reference_checkpoint_layout(tmp_path):
    """A checkpoint with the reference's exact field layout
    (src/fhs/facetHyperSARA.m:1165-1219: per data worker k, epsilon / v2 /
    proj / t_block / norm_res hold a cell per channel of a cell per block;
    param carries the init_* restart fields), as MATLAB writes it after a
    re-save with -v7 (the solver's own matfile(..., 'Writable', true) writes
    v7.3, which is refused here with a message)."""
    import scipy.io as sio
    rng = np.random.default_rng(4)
    Ny, Nx, L, K, Q = 6, 5, 2, 2, 2

    def cell(items):
        c = np.empty((len(items), 1), dtype=object)
        for i, it in enumerate(items):
            c[i, 0] = it
        return c
    blocks = [[3, 2], [4]]                            # per channel of a worker: block sizes
    v2 = cell([cell([cell([rng.normal(size=(m, 1)) + 1j * rng.normal(size=(m, 1)) for m in blocks[c]])
                     for c in range(len(blocks))]) for _ in range(K)])
    eps = cell([cell([cell([np.array([[1.5]]) for _ in blocks[c]]) for c in range(len(blocks))]) for _ in range(K)])
    m = {'param': {'pdfb_max_iter': 2000.0, 'reweighting_alpha': 1.0, 'init_reweight_step_count': 2.0,
                   'init_reweight_last_iter_step': 640.0, 'init_t_start': 641.0},
         'res': rng.normal(size=(Ny, Nx, L)), 'g': rng.normal(size=(Ny, Nx, L)), 'xsol': rng.random((Ny, Nx, L)),
         'epsilon': eps, 'v2': v2, 'proj': v2, 't_block': eps, 'norm_res': eps,
         'v0': cell([rng.normal(size=(12, L)) for _ in range(Q)]), 'v1': cell([rng.normal(size=(50, L))
                                                                                for _ in range(Q)]),
         'weights0': cell([np.ones((L, 1)) for _ in range(Q)]), 'weights1': cell([np.ones((50, 1))
                                                                                 for _ in range(Q)]),
         'end_iter': np.arange(641.0)[None, :], 't_facet': np.zeros((1, 641)), 't_data': np.zeros((1, 641))}
    p = tmp_path / 'fhs_cube_rw2.mat'
    sio.savemat(p, m, format='5')
    c = fio.load_checkpoint(p)
    assert set(c) == set(fio.CHECKPOINT_FIELDS) - {'rel_val'}
    assert np.array_equal(c['xsol'], m['xsol'])
    assert len(c['v2']) == K and len(c['v0']) == Q
    # worker k, channel 1, block 0: the nested cells survive
    w = c['v2'][1]
    assert np.allclose(np.asarray(w.ravel()[1].ravel()[0]).ravel(), v2[1, 0][1, 0][0, 0].ravel())
    assert int(np.asarray(c['param']['init_t_start'][0, 0]).squeeze()) == 641
    assert c['end_iter'].size == 641
