"""On-disk formats of the reference around the iteration: per-channel data
files, checkpoints and FITS cubes.

- ``load_data_ch``: ``data_ch_<k>.mat`` with the fields the synthetic-data
  generator writes (``imaging/main_generate_data.m:345-364``: ``y0, y,
  l2bounds, sigma_noise, u, v, w, frequency, maxProjBaseline, nW``; real
  data adds ``nWimag``) and that the imaging driver reads
  (``imaging/imaging.m:378,456-463``).  MAT v5/v7 files are read with
  scipy; MAT v7.3 files are HDF5, which needs h5py (not part of this
  environment): they are detected and refused with an explicit message
  (re-save with ``save(..., '-v7')``).  ``save_data_ch`` writes the same
  fields (v5).
- ``half_spatial_bandwidth`` / ``uv_to_radians``: the pixel-size rule of
  ``imaging/imaging.m:373-386`` and the scaling of
  ``lib/operators/util_load_data_gen_measurement_operator.m:89-91``
  (``u * pi / halfSpatialBandwidth``, ``-v * pi / halfSpatialBandwidth``).
- ``save_checkpoint`` / ``load_checkpoint``: ``<name>_rw<flag>.mat`` with the
  fields of ``src/fhs/facetHyperSARA.m:1165-1205`` (``src/hs/hyperSARA.m``
  the same names): ``param, res, g, xsol, epsilon, v2, proj, t_block,
  norm_res, v0, v1, weights0, weights1, end_iter, t_facet, t_data,
  rel_val`` (cells as object arrays, nested cells kept as nested object
  arrays).  The exchange works one way as it stands: the reference reads the
  v5 files written here, but it writes its own checkpoints through
  ``matfile(..., 'Writable', true)`` (``:1165``), which makes MAT v7.3
  (HDF5) files; those must be re-saved with ``-v7`` before they load here.
- ``fitswrite`` / ``fitsread``: the primary-HDU FITS files of the model and
  residual cubes (``..._FINAL_WB_MODEL.fits``, ``..._FINAL_WB_RESIDUAL.fits``,
  ``src/fhs/facetHyperSARA.m:1218-1219``) and of initial images
  (``imaging/imaging.m:886``), written as MATLAB's ``fitswrite`` lays out an
  Ny x Nx x L array: NAXIS1 = Ny (fastest), NAXIS2 = Nx, NAXIS3 = L,
  big-endian IEEE floats (BITPIX -64 for double, -32 for single), header and
  data padded to 2880-byte records (FITS standard 4.0).
"""
import numpy as np

FITS_BLOCK = 2880
ARCSEC_PER_RAD = (180.0 / np.pi) * 3600.0


# ------------------------------------------------------------------ MAT --
def _is_hdf5_mat(path):
    with open(path, 'rb') as fh:
        head = fh.read(128)
    return head[:19] == b'MATLAB 7.3 MAT-file' or head[:8] == b'\x89HDF\r\n\x1a\n'


DATA_FIELDS = ('y0', 'y', 'l2bounds', 'sigma_noise', 'u', 'v', 'w', 'frequency', 'maxProjBaseline', 'nW',
               'nWimag')


def load_data_ch(path, fields=None):
    """Fields of a ``data_ch_<k>.mat`` file as numpy arrays (vectors
    flattened, scalars as Python floats); missing fields are absent."""
    import scipy.io as sio
    if _is_hdf5_mat(path):
        raise ValueError(f'{path}: MAT v7.3 (HDF5) file; reading it needs h5py, which is not available here. '
                         f"Re-save it in MATLAB with save(name, '-struct', 's', '-v7').")
    want = list(fields or DATA_FIELDS)
    raw = sio.loadmat(path, variable_names=want)
    out = {}
    for k in want:
        if k not in raw:
            continue
        a = np.asarray(raw[k])
        out[k] = a.item() if a.size == 1 and k in ('frequency', 'maxProjBaseline', 'sigma_noise') else a.ravel(
            order='F')
    return out


def save_data_ch(path, **fields):
    """Write a ``data_ch_<k>.mat`` (MAT v5) with the reference's field names;
    vectors are stored as MATLAB column vectors."""
    import scipy.io as sio
    unknown = set(fields) - set(DATA_FIELDS)
    if unknown:
        raise ValueError(f'save_data_ch: unknown fields {sorted(unknown)}')
    mdict = {}
    for k, v in fields.items():
        a = np.asarray(v)
        mdict[k] = a.reshape(-1, 1) if a.ndim == 1 else a
    sio.savemat(path, mdict, format='5', do_compression=False)


def half_spatial_bandwidth(max_proj_baseline=None, superresolution=None, pixel_size=None):
    """``imaging/imaging.m:373-386``: with no pixel size (the 'nominal'
    resolution), pixelSize = (180/pi) 3600 / (superresolution * 2 *
    maxProjBaseline) arcsec; halfSpatialBandwidth = (180/pi) 3600 /
    pixelSize / 2."""
    if pixel_size is None:
        if max_proj_baseline is None or superresolution is None:
            raise ValueError('half_spatial_bandwidth: give pixel_size, or max_proj_baseline and superresolution')
        pixel_size = ARCSEC_PER_RAD / (superresolution * 2.0 * max_proj_baseline)
    return ARCSEC_PER_RAD / pixel_size / 2.0


def uv_to_radians(u, v, half_bandwidth):
    """``util_load_data_gen_measurement_operator.m:89-91``: the NUFFT
    frequencies (u along x, v along y, in radians per pixel)."""
    return (np.asarray(u, dtype=np.float64) * np.pi / half_bandwidth,
            -np.asarray(v, dtype=np.float64) * np.pi / half_bandwidth)


# ----------------------------------------------------------- checkpoint --
CHECKPOINT_FIELDS = ('param', 'res', 'g', 'xsol', 'epsilon', 'v2', 'proj', 't_block', 'norm_res', 'v0', 'v1',
                     'weights0', 'weights1', 'end_iter', 't_facet', 't_data', 'rel_val')


def _cell(items):
    c = np.empty((len(items), 1), dtype=object)
    for i, it in enumerate(items):
        c[i, 0] = np.asarray(it)
    return c


def save_checkpoint(path, xsol, res, g, epsilon, v2, proj, t_block, norm_res, v0, v1, weights0, weights1,
                    param=None, end_iter=None, t_facet=None, t_data=None, rel_val=None):
    """``<name>_rw<flag>.mat`` (MAT v5).  Cubes (Ny, Nx, L); per-data-worker
    lists (epsilon, v2, proj, t_block, norm_res: one entry per channel group)
    and per-facet lists (v0, v1, weights0, weights1) become K x 1 / Q x 1
    cells."""
    import scipy.io as sio
    m = {'xsol': np.asarray(xsol, dtype=np.float64), 'res': np.asarray(res, dtype=np.float64),
         'g': np.asarray(g, dtype=np.float64),
         'epsilon': _cell(epsilon), 'v2': _cell(v2), 'proj': _cell(proj), 't_block': _cell(t_block),
         'norm_res': _cell(norm_res), 'v0': _cell(v0), 'v1': _cell(v1), 'weights0': _cell(weights0),
         'weights1': _cell(weights1)}
    if param is not None:
        m['param'] = {k: v for k, v in param.items() if v is not None}
    for k, v in (('end_iter', end_iter), ('t_facet', t_facet), ('t_data', t_data), ('rel_val', rel_val)):
        if v is not None:
            m[k] = np.asarray(v, dtype=np.float64)
    sio.savemat(path, m, format='5', do_compression=False)


def load_checkpoint(path):
    """Fields of a checkpoint written by ``save_checkpoint`` (or by MATLAB in
    MAT v5/v7): cubes as arrays, cells as lists."""
    import scipy.io as sio
    if _is_hdf5_mat(path):
        raise ValueError(f'{path}: MAT v7.3 (HDF5) file; reading it needs h5py, which is not available here')
    raw = sio.loadmat(path)
    out = {}
    for k in CHECKPOINT_FIELDS:
        if k not in raw:
            continue
        a = raw[k]
        if a.dtype == object and k != 'param':
            out[k] = [np.asarray(x).squeeze() for x in a.ravel(order='F')]
        else:
            out[k] = a
    return out


# ------------------------------------------------------------------ FITS --
def _card(key, value=None, comment=None):
    if value is None:
        s = f'{key:<8}'
    else:
        if isinstance(value, bool):
            v = 'T' if value else 'F'
        elif isinstance(value, (int, np.integer)):
            v = str(int(value))
        else:
            v = repr(float(value)).upper()
        s = f'{key:<8}= {v:>20}'
        if comment:
            s += f' / {comment}'
    if len(s) > 80:
        raise ValueError(f'FITS card too long: {s!r}')
    return s.ljust(80)


def fitswrite(data, path, dtype=np.float64):
    """Write ``data`` (an Ny x Nx [x L] array) as the primary HDU of a FITS
    file, laid out as MATLAB's ``fitswrite``: NAXIS1 = size(data, 1)."""
    a = np.asarray(data, dtype=dtype)
    bitpix = {np.dtype(np.float64): -64, np.dtype(np.float32): -32}.get(a.dtype)
    if bitpix is None:
        raise ValueError('fitswrite: float64 or float32 data only')
    cards = [_card('SIMPLE', True, 'conforms to FITS standard'), _card('BITPIX', bitpix),
             _card('NAXIS', a.ndim)]
    for i, n in enumerate(a.shape):
        cards.append(_card(f'NAXIS{i + 1}', n))
    cards.append(_card('END'))
    header = ''.join(cards).encode('ascii')
    header += b' ' * (-len(header) % FITS_BLOCK)
    body = np.asfortranarray(a).astype(a.dtype.newbyteorder('>')).tobytes(order='F')
    body += b'\0' * (-len(body) % FITS_BLOCK)
    with open(path, 'wb') as fh:
        fh.write(header)
        fh.write(body)


def fitsread(path):
    """The primary HDU's data of a FITS file (BITPIX -64/-32/8/16/32/64,
    BSCALE/BZERO applied) as an array indexed (NAXIS1, NAXIS2, ...), like
    MATLAB's ``fitsread``."""
    with open(path, 'rb') as fh:
        raw = fh.read()
    keys, pos = {}, 0
    while True:
        if pos + 80 > len(raw):
            raise ValueError(f'{path}: FITS header without END')
        card = raw[pos:pos + 80].decode('ascii')
        pos += 80
        key = card[:8].strip()
        if key == 'END':
            break
        if card[8:10] == '= ':
            val = card[10:].split('/')[0].strip()
            keys[key] = val
    if keys.get('SIMPLE') != 'T':
        raise ValueError(f'{path}: not a FITS primary header (SIMPLE != T)')
    pos += -pos % FITS_BLOCK
    bitpix = int(keys['BITPIX'])
    naxis = int(keys['NAXIS'])
    shape = tuple(int(keys[f'NAXIS{i + 1}']) for i in range(naxis))
    dt = {-64: '>f8', -32: '>f4', 8: 'u1', 16: '>i2', 32: '>i4', 64: '>i8'}[bitpix]
    n = int(np.prod(shape)) if shape else 0
    a = np.frombuffer(raw, dtype=dt, count=n, offset=pos).reshape(shape, order='F')
    a = a.astype(a.dtype.newbyteorder('='))
    bscale, bzero = float(keys.get('BSCALE', 1.0)), float(keys.get('BZERO', 0.0))
    if bscale != 1.0 or bzero != 0.0:
        a = a * bscale + bzero
    return a
