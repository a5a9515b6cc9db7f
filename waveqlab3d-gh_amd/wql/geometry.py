"""Block geometry: node coordinates X[3][ns][nr][nq] of one block.

Host-side initialisation (runs once per run, not on the time-stepping path).
The coordinates are the input of the metric computation done on the device.

* ``cartesian`` restates cartesian_grid_3d (grid.f90:213-240).
* ``curvilinear`` restates the transfinite-interpolation generator
  Curve_Grid_3D3 (grid.f90:471-1907) for the cases the hot path needs: the
  'analytical_test_problem' fault profile (grid.f90:775-831), the
  'analytical_tpv28' hills, the 'analytical_tpv10' 60-degree and
  'analytical_tpv36' 15-degree dipping faults (grid.f90:625-773), the
  'read_from_memomry_fractal' fault surface read from a text file
  (grid.f90:832-925, read_2darray 2457-2481), plus
  caller-supplied fault-surface and free-surface perturbation fields (the
  seeded topography / rough-fault inputs of the benchmark workloads, which the
  reference reads from files through the same TFI; grid.f90:1349-1430).
  Face TFI: Gen_Surf_Mesh (grid.f90:2418-2455); volume TFI: Gen3_Curve_Mesh
  (grid.f90:2290-2416).
"""
import numpy as np

PI = 3.141592653589793


def cartesian(n, a, b):
    """X for a uniform block with n=(nq,nr,ns) nodes spanning [a, b]."""
    nq, nr, ns = n
    X = np.empty((3, ns, nr, nq))
    for d, m in enumerate((nq, nr, ns)):
        h = (b[d] - a[d]) / float(m - 1)
        c = a[d] + np.arange(m, dtype=np.float64) * h
        shape = [1, 1, 1]
        shape[2 - d] = m
        X[d] = np.broadcast_to(c.reshape(shape), (ns, nr, nq))
    return X


def _surf(mx, my, XL, XR, XB, XT):
    """Gen_Surf_Mesh: 2-D TFI; result indexed [i, j] with i < mx, j < my."""
    hq = 1.0 / float(mx - 1)
    hr = 1.0 / float(my - 1)
    q = (np.arange(mx, dtype=np.float64) * hq)[:, None]
    r = (np.arange(my, dtype=np.float64) * hr)[None, :]
    XL = np.asarray(XL)[None, :]
    XR = np.asarray(XR)[None, :]
    XBi = np.asarray(XB)[:, None]
    XTi = np.asarray(XT)[:, None]
    return ((1.0 - q) * XL + q * XR + (1.0 - r) * XBi + r * XTi
            - (1.0 - q) * (1.0 - r) * XL[:, :1] - q * (1.0 - r) * XR[:, :1]
            - r * (1.0 - q) * XT[0] - (r * q) * XT[mx - 1])


def _tfi3(n, Xleft, Xright, Xtop, Xbottom, Xfront, Xback):
    """Gen3_Curve_Mesh for one coordinate component -> array [ns][nr][nq].

    Face arrays: left/right [nr][ns] (j,k), top/bottom [nq][ns] (i,k),
    front/back [nq][nr] (i,j)."""
    nq, nr, ns = n
    hq, hr, hs = 1.0 / float(nq - 1), 1.0 / float(nr - 1), 1.0 / float(ns - 1)
    q = (np.arange(nq, dtype=np.float64) * hq)[None, None, :]
    r = (np.arange(nr, dtype=np.float64) * hr)[None, :, None]
    s = (np.arange(ns, dtype=np.float64) * hs)[:, None, None]
    L = Xleft.T[:, :, None]       # [k][j][1]
    R = Xright.T[:, :, None]
    T = Xtop.T[:, None, :]        # [k][1][i]
    B = Xbottom.T[:, None, :]
    Fr = Xfront.T[None, :, :]     # [1][j][i]
    Bk = Xback.T[None, :, :]
    U = (1.0 - q) * L + q * R
    V = (1.0 - r) * T + r * B
    W = (1.0 - s) * Fr + s * Bk
    UW = ((1.0 - q) * ((1.0 - r) * Xleft[0, :][:, None, None] + r * Xleft[nr - 1, :][:, None, None])
          + q * ((1.0 - r) * Xright[0, :][:, None, None] + r * Xright[nr - 1, :][:, None, None]))
    UV = ((1.0 - r) * ((1.0 - s) * Xtop[:, 0][None, None, :] + s * Xtop[:, ns - 1][None, None, :])
          + r * ((1.0 - s) * Xbottom[:, 0][None, None, :] + s * Xbottom[:, ns - 1][None, None, :]))
    VW = ((1.0 - s) * ((1.0 - q) * Xfront[0, :][None, :, None] + q * Xfront[nq - 1, :][None, :, None])
          + s * ((1.0 - q) * Xback[0, :][None, :, None] + q * Xback[nq - 1, :][None, :, None]))
    UVW1 = (1.0 - q) * ((1.0 - r) * ((1.0 - s) * Xtop[0, 0] + s * Xtop[0, ns - 1])
                        + r * ((1.0 - s) * Xbottom[0, 0] + s * Xbottom[0, ns - 1]))
    UVW2 = q * ((1.0 - r) * ((1.0 - s) * Xtop[nq - 1, 0] + s * Xtop[nq - 1, ns - 1])
                + r * ((1.0 - s) * Xbottom[nq - 1, 0] + s * Xbottom[nq - 1, ns - 1]))
    return U + V + W - UV - UW - VW + UVW1 + UVW2


# fault-surface profiles of Curve_Grid_3D3 given by a formula: x offset f(y)
# or f(y, z) of the q faces, weighted by rc (left face) and lc (right face)
ANALYTICAL_PROFILES = ('analytical_tpv28', 'analytical_tpv10', 'analytical_tpv36')


def profile_offset(profile, Y, Z, ay, by, use):
    """f of an analytical fault profile at face nodes (Y, Z); `use` False
    gives 0 (the dipping profiles evaluate f only on a face whose weight is
    positive)"""
    f = np.zeros_like(Y)
    if profile == 'analytical_tpv28':  # grid.f90:627-656: two cosine hills
        r1 = np.sqrt((Z + 10.5) ** 2 + (Y - 7.5) ** 2)
        r2 = np.sqrt((Z - 10.5) ** 2 + (Y - 7.5) ** 2)
        f = np.where(r1 <= 3.0, -0.3 * (1.0 + np.cos(PI * r1 / 3.0)), f)
        f = np.where(r2 <= 3.0, -0.3 * (1.0 + np.cos(PI * r2 / 3.0)), f)
    elif profile == 'analytical_tpv10' and use:  # 658-713: 60 degrees, rounded below 15 km
        ang = 60.0 / 180.0 * PI
        deep = (1.0 / np.tan(ang)) * (15.0 - 0.5 * (ay + by)) + \
            0.5 * np.arctan(4.0 * (Y - 15.0)) * np.exp(-5.0 * (Y - 15.0))
        f = np.where(Y <= 15.0, (1.0 / np.tan(ang)) * (Y - 0.5 * (ay + by)), deep)
    elif profile == 'analytical_tpv36' and use:  # 715-773: 15 degrees to 8 km depth
        ang = 15.0 / 180.0 * PI
        deep = 1.0 / np.tan(ang) * 8.0 - 15.0 + 0.25 * np.arctan(4.0 * (Y - 8.0)) * np.exp(-0.0 * (Y - 8.0))
        f = np.where(Y <= 8.0, 1.0 / np.tan(ang) * Y - 15.0, deep)
    return f


def read_fault_profile(path, my, mz):
    """x offsets [my][mz] of a 'read_from_memomry_fractal' fault surface: a
    text file of my*mz rows 'x y z', y fastest (read_2darray and
    reshape_2darray, grid.f90:2457-2505); only x is used (grid.f90:832-925)"""
    d = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if d.shape[0] < my * mz or d.shape[1] < 3:
        raise ValueError(f'fault profile {path}: {d.shape[0]} rows of {d.shape[1]} values, '
                         f'{my * mz} rows of 3 needed')
    return d[:my * mz, 0].reshape(mz, my).T.copy()


def curvilinear(n, a, b, lc=0.0, rc=0.0, profile='analytical_test_problem',
                fault_left=None, fault_right=None, top=None, profile_path=None):
    """Curvilinear block with n=(nq,nr,ns) nodes over [a, b].

    profile 'analytical_test_problem' adds the Gaussian-plus-sinusoid bump of
    grid.f90:812-827 to the q-faces, weighted by rc (left face) and lc (right
    face); 'analytical_tpv28', 'analytical_tpv10' and 'analytical_tpv36' add
    their formulas the same way, 'read_from_memomry_fractal' the x column of
    the file `profile_path`.  Any other profile keeps the q-faces planar (as
    the reference does for names it does not know); then ``fault_left`` /
    ``fault_right`` ([nr][ns] arrays of x offsets) perturb the q-faces and
    ``top`` ([nq][ns] array of y offsets) perturbs the r = 1 face.
    """
    mx, my, mz = n
    ax, ay, az = a
    bx, by, bz = b
    hy = (by - ay) / float(my - 1)
    hz = (bz - az) / float(mz - 1)
    jy = ay + np.arange(my, dtype=np.float64) * hy
    kz = az + np.arange(mz, dtype=np.float64) * hz
    # x-faces (grid.f90:779-811)
    Yleft = _surf(my, mz, np.full(mz, ay), np.full(mz, by), jy, jy)
    Zleft = _surf(my, mz, kz, kz, np.full(my, az), np.full(my, bz))
    Yright, Zright = Yleft.copy(), Zleft.copy()
    if profile == 'analytical_test_problem':
        Xleft = (ax + 0.1 * rc * np.exp(-((1.0 * Yleft - 0.5 * (ay + by)) ** 2
                                          + 1.0 * (Zleft - 0.5 * (az + bz)) ** 2) / 0.025)
                 + 0.1 * rc * np.sin(2.0 / (by - ay) * PI * Yleft) * np.sin(2.0 / (bz - az) * PI * Zleft))
        Xright = (bx + 0.1 * lc * np.exp(-(1.0 * (Yright - 0.5 * (ay + by)) ** 2
                                           + 1.0 * (Zright - 0.5 * (az + bz)) ** 2) / 0.025)
                  + 0.1 * lc * np.sin(2.0 / (by - ay) * PI * Yright) * np.sin(2.0 / (bz - az) * PI * Zright))
    elif profile in ANALYTICAL_PROFILES:
        # these branches set the face nodes' y and z directly (grid.f90:636-639)
        Yleft = np.broadcast_to(jy[:, None], (my, mz)).copy()
        Zleft = np.broadcast_to(kz[None, :], (my, mz)).copy()
        Yright, Zright = Yleft.copy(), Zleft.copy()
        # one f for both faces: the dipping profiles evaluate it when either
        # weight is positive (from Yleft or Yright, which are equal)
        f = profile_offset(profile, Yleft, Zleft, ay, by, rc > 0.0 or lc > 0.0)
        Xleft = ax + rc * f
        Xright = bx + lc * f
    elif profile == 'read_from_memomry_fractal':
        X1 = read_fault_profile(profile_path, my, mz)
        Xleft = np.full((my, mz), float(ax))
        Xright = np.full((my, mz), float(bx))
        if rc != 0.0:
            Xleft = Xleft + rc * X1
        if lc != 0.0:
            Xright = Xright + lc * X1
    else:
        Xleft = np.full((my, mz), float(ax))
        Xright = np.full((my, mz), float(bx))
        if fault_left is not None:
            Xleft = Xleft + fault_left
        if fault_right is not None:
            Xright = Xright + fault_right

    def edge(lo, hi, m):
        h = (hi - lo) / float(m - 1)
        return lo + np.arange(m, dtype=np.float64) * h

    # z-faces (grid.f90:930-1041)
    faces = {}
    for tag, k in (('front', 0), ('back', mz - 1)):
        XL, XR = Xleft[:, k], Xright[:, k]
        YL, YR = Yleft[:, k], Yright[:, k]
        XB = edge(Xleft[0, k], Xright[0, k], mx)
        XT = edge(Xleft[my - 1, k], Xright[my - 1, k], mx)
        YB = edge(Yleft[0, k], Yright[0, k], mx)
        YT = edge(Yleft[my - 1, k], Yright[my - 1, k], mx)
        faces['X' + tag] = _surf(mx, my, XL, XR, XB, XT)
        faces['Y' + tag] = _surf(mx, my, YL, YR, YB, YT)
        faces['Z' + tag] = np.full((mx, my), az if k == 0 else bz)
    # y-faces (grid.f90:1044-1158)
    for tag, j in (('top', 0), ('bottom', my - 1)):
        ZT = faces['Zback'][:, j]
        ZB = faces['Zfront'][:, j]
        XB = edge(Xleft[j, 0], Xright[j, 0], mx)
        XT = edge(Xleft[j, mz - 1], Xright[j, mz - 1], mx)
        XL, XR = Xleft[j, :], Xright[j, :]
        ZL, ZR = Zleft[j, :], Zright[j, :]
        faces['X' + tag] = _surf(mx, mz, XL, XR, XB, XT)
        faces['Z' + tag] = _surf(mx, mz, ZL, ZR, ZB, ZT)
        faces['Y' + tag] = np.full((mx, mz), ay if j == 0 else by)
    if top is not None:
        faces['Ytop'] = faces['Ytop'] + top
    X = np.empty((3, mz, my, mx))
    X[0] = _tfi3(n, Xleft, Xright, faces['Xtop'], faces['Xbottom'], faces['Xfront'], faces['Xback'])
    X[1] = _tfi3(n, Yleft, Yright, faces['Ytop'], faces['Ybottom'], faces['Yfront'], faces['Yback'])
    X[2] = _tfi3(n, Zleft, Zright, faces['Ztop'], faces['Zbottom'], faces['Zfront'], faces['Zback'])
    return X


def read_topography(path, mx, mz):
    """XT, YT, ZT [mx][mz] of a 'read_topo_from_memory' file: mx*mz rows
    'x y z', z fastest (read_2darray and reshape_2darray to (mz, mx),
    grid.f90:1394-1420)"""
    d = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if d.shape[0] < mx * mz or d.shape[1] < 3:
        raise ValueError(f'topography {path}: {d.shape[0]} rows of {d.shape[1]} values, {mx * mz} rows of 3 needed')
    return [d[:mx * mz, c].reshape(mx, mz) for c in range(3)]


def topography_faces(n, a, b, lc=0.0, rc=0.0, profile='no_path', profile_path=None, faultsize=None,
                     topography='read_topo_from_memory', topography_path=None, topo=1.0):
    """The six faces of Curve_Grid_3D3's topography branch (use_topography,
    grid.f90:1349-1860): the top face r = 0 from the topography, the q faces
    interpolated between it and y = b_y (with the x column, y and z of a
    'read_from_memomry_fractal' file of faultsize rows on a weighted face),
    the s faces and the bottom face from their edges.  Returns
    {'X': (left, right, top, bottom, front, back), 'Y': ..., 'Z': ...} in
    _tfi3's layouts."""
    mx, my, mz = n
    ax, ay, az = a
    bx, by, bz = b
    Ybottom = np.full((mx, mz), float(by))
    if topography == 'analytical_topography':  # 1364-1392 (both radii from (10, 10))
        hx = (bx - ax) / float(mx - 1)
        hz = (bz - az) / float(mz - 1)
        Xtop = np.broadcast_to((ax + np.arange(mx, dtype=np.float64) * hx)[:, None], (mx, mz)).copy()
        Ztop = np.broadcast_to((az + np.arange(mz, dtype=np.float64) * hz)[None, :], (mx, mz)).copy()
        r1 = np.sqrt((Xtop - 10.0) ** 2 + (Ztop - 10.0) ** 2)
        f = np.where(r1 <= 3.0, -0.3 * (1.0 + np.cos(PI * r1 / 3.0)), 0.0)
        Ytop = ay - topo * 1.5 * f
    elif topography == 'read_topo_from_memory':  # 1394-1420
        XT, YT, ZT = read_topography(topography_path, mx, mz)
        Xtop, Ytop, Ztop = XT, ay + topo * YT, ZT
    else:
        raise ValueError(f"specify topography profile (got {topography!r})")
    hy, hz = 1.0 / float(my - 1), 1.0 / float(mz - 1)
    r = np.arange(my, dtype=np.float64) * hy
    sz = np.arange(mz, dtype=np.float64) * hz
    zc = az + (bz - az) * sz

    def qface(i):  # 1422-1480: the q face at column i of the top face
        YL, YR = Ytop[i, :], np.full(mz, float(by))
        YB = Ytop[i, 0] + (by - Ytop[i, 0]) * r
        YT_ = Ytop[i, mz - 1] + (by - Ytop[i, mz - 1]) * r
        return _surf(my, mz, YL, YR, YB, YT_), _surf(my, mz, zc, zc, np.full(my, float(az)), np.full(my, float(bz)))
    Yleft, Zleft = qface(0)
    Yright, Zright = qface(mx - 1)
    Xleft = np.full((my, mz), float(ax))
    Xright = np.full((my, mz), float(bx))
    if profile == 'read_from_memomry_fractal':  # 1561-1610
        ny, nz = faultsize if faultsize else (my, mz)
        d = np.loadtxt(profile_path, dtype=np.float64, ndmin=2)
        if d.shape[0] < ny * nz or d.shape[1] < 3:
            raise ValueError(f'fault profile {profile_path}: {ny * nz} rows of 3 needed')
        X1, Y1, Z1 = [d[:ny * nz, c].reshape(nz, ny).T[:my, :mz] for c in range(3)]
        if rc != 0.0:
            Xleft, Yleft, Zleft = ax + X1, Y1.copy(), Z1.copy()
        if lc != 0.0:
            Xright, Yright, Zright = bx + X1, Y1.copy(), Z1.copy()
    elif profile != 'no_path':
        raise ValueError(f"specify fault profile (the topography branch takes 'no_path' or "
                         f"'read_from_memomry_fractal', got {profile!r})")
    hx = 1.0 / float(mx - 1)
    q = np.arange(mx, dtype=np.float64) * hx

    def sface(k):  # 1700-1745: the s face at plane k
        XT_ = Xleft[my - 1, k] + (Xright[my - 1, k] - Xleft[my - 1, k]) * q
        X = _surf(mx, my, Xleft[:, k], Xright[:, k], Xtop[:, k], XT_)
        Y = _surf(mx, my, Yleft[:, k], Yright[:, k], Ytop[:, k], Ybottom[:, k])
        return X, Y
    Xfront, Yfront = sface(0)
    Xback, Yback = sface(mz - 1)
    Zfront = np.full((mx, my), float(az))
    Zback = np.full((mx, my), float(bz))
    # bottom face (1747-1762)
    Xbottom = _surf(mx, mz, Xleft[my - 1, :], Xright[my - 1, :], Xfront[:, my - 1], Xback[:, my - 1])
    Zbottom = _surf(mx, mz, Zleft[my - 1, :], Zright[my - 1, :], Zfront[:, my - 1], Zback[:, my - 1])
    return {'X': (Xleft, Xright, Xtop, Xbottom, Xfront, Xback),
            'Y': (Yleft, Yright, Ytop, Ybottom, Yfront, Yback),
            'Z': (Zleft, Zright, Ztop, Zbottom, Zfront, Zback)}


def curvilinear_topography(n, a, b, **kw):
    """X[3][ns][nr][nq] of Curve_Grid_3D3's topography branch (faces of
    topography_faces, volume Gen3_Curve_Mesh)"""
    faces = topography_faces(n, a, b, **kw)
    mx, my, mz = n
    X = np.empty((3, mz, my, mx))
    for c, key in enumerate('XYZ'):
        X[c] = _tfi3(n, *faces[key])
    return X

