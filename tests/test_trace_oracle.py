"""Pins of the oracle's restatement of TriMesh::computeStandardStretch (TriMesh.cpp:1585-1652), the info.txt
distortion metrics, by closed forms: an isometry gives (L2, Linf, shear, compress) = (1, 1, 0, 1); a similarity
(rotation and uniform scale) gives the same after the reference's area normalisation; the axis scaling (x, 2y) of
a flat mesh has deformation-gradient singular values 1 and 1/2 in every triangle, so L2 = sqrt(1.25 / 2 * 2),
Linf = sqrt(2) and compress = sqrt(2) / 2 after normalisation by sqrt(UV area / rest area) = sqrt(2)."""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd.meshgen import height_field_disc


def flat_disc(m=9, seed=3):
    V, F, UV = height_field_disc(m, seed=seed)
    V = V.copy()
    V[:, 2] = 0.0
    return V, F, UV


def test_stretch_isometry_and_similarity():
    V, F, UV = flat_disc()
    np.testing.assert_allclose(O.standard_stretch(V, F, UV), (1.0, 1.0, 0.0, 1.0), atol=1e-12)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    np.testing.assert_allclose(O.standard_stretch(V, F, 3.5 * UV @ R.T + 2.0), (1.0, 1.0, 0.0, 1.0), atol=1e-12)
    # the rest mesh moved rigidly in 3D changes nothing either
    Rz = np.array([[1, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])
    np.testing.assert_allclose(O.standard_stretch(V @ Rz.T + 5.0, F, UV), (1.0, 1.0, 0.0, 1.0), atol=1e-12)


def test_stretch_axis_scaling_closed_form():
    V, F, UV = flat_disc()
    l2, linf, shear, comp = O.standard_stretch(V, F, UV * np.array([1.0, 2.0]))
    assert l2 == pytest.approx(np.sqrt(1.25), rel=1e-12)
    assert linf == pytest.approx(np.sqrt(2.0), rel=1e-12)
    assert comp == pytest.approx(np.sqrt(2.0) / 2.0, rel=1e-12)
    assert shear > 0.0
