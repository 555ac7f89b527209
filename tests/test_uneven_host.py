"""Host side of the packed uneven instances (nexd_amd.uneven), no GPU."""
import pytest


def test_pack_refuses_pml_instances():
    """Packed unions share one mesh; the PML update of the reference reads
    the mesh's last element, so PML instances are refused."""
    from nexd_amd import batch, problem, uneven
    p = problem.layered_box(nx=2, ny=2, nz=2, order=2, nt=4, pml_cells=1, nrec_side=1, timeint=2,
                            attenuation=False)
    with pytest.raises(ValueError, match="PML"):
        uneven.pack([batch.ShotBatch(p, [p], True)])
