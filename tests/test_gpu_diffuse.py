"""The day's diffuse patch sums on the matrix cores (sw_day_diffuse). For every pixel and daytime
hour the device result must be the exact integer product of the pixel's sky | vegetation-only bits
with the day's diffuse table, rounded once to FP32 (the same bits as that product formed here in
int64), for every tile width; and it must lie within the table's quantisation bound of the FP64
diffsh-weighted patch sums (umep_solweig.py:148-154: diffsh 1 on sky patches, the hour's
vegetation diffsh on vegetation-only patches, 0 on building patches). Night slots are not written."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _mask_bits(vis, P):
    """[P][320] 0/1 of the 320 table positions: sky bits (vsh & vveg) at 0..159, vegetation-only
    bits (vsh & ~vveg) at 160..319, from the visibility words [3][5][P]."""
    vsh, vveg = vis[0], vis[1]
    words = np.concatenate([vsh & vveg, vsh & ~vveg])                   # [10][P]
    bits = (words[:, :, None] >> np.arange(32, dtype=np.uint32)) & np.uint32(1)
    return bits.transpose(1, 0, 2).reshape(P, 320).astype(np.int64)


def _table_integers(table, n_hours):
    from solweig_amd import sky
    digits, scale = sky.diffuse_digits(table, n_hours)
    d = digits.astype(np.int64)
    return d[0] + d[1] * 128 + d[2] * 128 ** 2 + d[3] * 128 ** 3, scale


@pytest.mark.parametrize('tile,n_hours', [(16, 24), (32, 24), (128, 24), (16, 13)])
def test_day_diffuse_sums_are_exact_products(rt, tile, n_hours):
    """13 hours: two 8-hour blocks of table rows, the second reaching past the day."""
    from solweig_amd import batched, synthetic, sun as psun
    doms = [synthetic.small_domain(seed=80 + i, size=n, latitude=lat, maps=(i % 2 == 0))
            for i, (n, lat) in enumerate([(45, 12.0), (31, 52.0), (38, 30.0)])]
    suns = [{k: v[:n_hours] for k, v in psun.met_sun(d.met, d.location, d.utc).items()} for d in doms]
    eng = batched.Engine(doms, n_hours=n_hours)
    eng.set_tiles(tile)
    tabs = eng.tables(sun=suns, met=[d.met[:n_hours] for d in doms])
    eng.load(tabs)
    for b in eng.buf:                                 # all-ones bits: a NaN left where nothing is written
        rt.call('sw_memset', b['diffuse'].ptr, 0xff, b['diffuse'].nbytes, eng.stream.handle)
    print(f'tile {tile}, {n_hours} h: sw_day_diffuse over {len(doms)} domains', flush=True)
    eng.diffuse()
    eng.stream.sync()
    n_day = 0
    for k, (pdom, t, b) in enumerate(zip(eng.domains, tabs, eng.buf)):
        P = pdom.P
        got = b['diffuse'].to_host(eng.stream)                          # [hours][P][2]
        bits = _mask_bits(pdom.dev['vis'].to_host(eng.stream).reshape(3, 5, P), P)
        q, scale = _table_integers(t.diffuse, n_hours)
        exact = bits @ q.T                                              # [P][rows], exact int64
        for h in range(n_hours):
            if not t.hours[h, rt.H['DAY']]:
                assert np.all(got[h].view(np.uint32) == 0xffffffff), (k, h)
                continue
            n_day += 1
            f = t.ptab_f[int(t.hours[h, rt.H['PATCH_TABLE']])].astype(np.float64)
            dveg = float(np.float32(t.hours[h, rt.H['DIFF_VEG']]))
            for v, col in ((0, rt.PF['LV']), (1, rt.PF['KD'])):
                r = 2 * h + v
                want = (exact[:, r] * scale[r]).astype(np.float32)      # exact x 2^n, one rounding
                assert np.array_equal(got[h, :, v], want), (k, h, v)
                ref = bits[:, :153] @ f[:, col] + dveg * (bits[:, 160:313] @ f[:, col])
                # each of the at most 306 terms is within half a table unit (scale / 2), plus the
                # final rounding to FP32
                bound = 306 * scale[r] / 2 + np.abs(ref) * 2.0 ** -24
                assert np.all(np.abs(got[h, :, v] - ref) <= bound), (k, h, v)
    assert n_day > 0
