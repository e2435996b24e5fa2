"""Model state on disk, in the reference's format (models/core_model.py:374-449,
model_object.py:408-428, _model_methods.py:437-482,
group_model_object.py:325-365, param/base.py:134-163,
param/parameter.py:598-645, image/window_object.py:311-320).

The round trips follow the reference's own tests (tests/test_model.py:259-295,
tests/test_group_models.py:112-165): a saved model, loaded by filename into a
new model, carries the same parameter values, windows and knobs.
``reference_state_sersic.yaml`` is laid out key for key as the reference
writes a Sersic model with an image PSF; it loads into the model it
describes and saves back to the same keys.  No GPU is used here.
"""

import io
import json
import os

import numpy as np
import pytest
import torch
import yaml

import astrophot_amd as ap
from astrophot_amd.models import AstroPhot_Model

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_state_sersic.yaml")


def target(nx=50, ny=40, scale=0.8):
    return ap.image.Target_Image(data=np.zeros((ny, nx)), pixelscale=scale, zeropoint=22.5)


def sersic(name, tar, **kw):
    return AstroPhot_Model(name=name, model_type="sersic galaxy model", target=tar,
                           parameters={"center": [20, 20], "PA": 60 * np.pi / 180, "q": 0.5, "n": 2,
                                       "Re": 5, "Ie": 1}, **kw)


def strip_identity(state):
    if isinstance(state, dict):
        return {k: strip_identity(v) for k, v in state.items() if k not in ("identity", "target_identity")}
    if isinstance(state, list):
        return [strip_identity(v) for v in state]
    return state


@pytest.mark.parametrize("ext", ["yaml", "json"])
def test_sersic_save_load(tmp_path, ext):
    psf = ap.utils.gaussian_psf(1.5, 11, 0.8)
    m = sersic("test sersic", target(), psf=psf, psf_mode="full", sampling_tolerance=1e-3)
    m["q"].uncertainty = 0.03
    fn = str(tmp_path / f"model.{ext}")
    m.save(fn)
    m2 = AstroPhot_Model(name="load model", filename=fn)
    assert type(m2) is type(m) and m2.name == "load model"
    for p in ("center", "q", "PA", "n", "Re", "Ie"):
        assert torch.equal(m2[p].value, m[p].value), p
        assert m2[p].limits == m[p].limits and m2[p].cyclic == m[p].cyclic and m2[p].units == m[p].units
    assert float(m2["q"].uncertainty) == 0.03
    assert m2.window == m.window and m2.window.get_state() == m.window.get_state()
    assert m2.psf_mode == "full" and m2.sampling_tolerance == 1e-3
    assert m2.integrate_mode == m.integrate_mode
    assert torch.equal(m2.psf.data.cpu(), m.psf.data.cpu())
    assert m2.psf.window.get_state() == m.psf.window.get_state()
    assert strip_identity(m2.get_state()) == strip_identity(dict(m.get_state(), name="load model"))


def test_locked_and_free_parameters_survive():
    m = sersic("s", target())
    m["n"].locked = True
    m2 = AstroPhot_Model(name="s2", filename=m.get_state())
    assert m2["n"].locked
    assert m2.parameters.size == m.parameters.size == 6
    assert torch.equal(m2.parameters.vector_values(), m.parameters.vector_values())
    assert list(m2.parameters.vector_names()) == list(m.parameters.vector_names())


def test_group_save_load(tmp_path):
    tar = target(51, 51, 1.0)
    mod1 = sersic("base model 1", tar, window=[[5, 40], [3, 45]])
    mod2 = sersic("base model 2", tar)
    mod2["center"].value = [30.0, 28.0]
    mod2["center"].locked = True
    smod = AstroPhot_Model(name="group model", model_type="group model", models=[mod1, mod2], target=tar)
    fn = str(tmp_path / "group.yaml")
    smod.save(fn)
    new = AstroPhot_Model(name="group model", filename=fn)
    assert type(new).__name__ == "Group_Model"
    assert list(new.models) == list(smod.models)
    assert new.parameters.size == smod.parameters.size == 12
    assert torch.equal(new.parameters.vector_values(), smod.parameters.vector_values())
    for a, b in zip(new.models.values(), smod.models.values()):
        assert a.window.get_state() == b.window.get_state()
        # every sub-model reads its values from the group's tree
        assert a.parameters is new.parameters[a.name]
    assert new.window.get_state() == smod.window.get_state()
    # the saved group state holds each sub-model without a parameter copy
    st = yaml.safe_load(open(fn))
    assert all("parameters" not in s for s in st["models"].values())
    assert [n["name"] for n in st["parameters"]["nodes"]] == ["base model 1", "base model 2"]


def test_group_load_onto_existing_models():
    tar = target(51, 51, 1.0)
    g = AstroPhot_Model(name="g", model_type="group model", models=[sersic("a", tar), sersic("b", tar)],
                        target=tar)
    g.models["b"]["Re"].value = 7.5
    state = g.get_state()
    h = AstroPhot_Model(name="g", model_type="group model", models=[sersic("a", tar), sersic("b", tar)],
                        target=tar)
    h.load(state)
    assert float(h.models["b"]["Re"].value) == 7.5
    assert torch.equal(h.parameters.vector_values(), g.parameters.vector_values())


def test_reference_format_file_loads_and_saves_back(tmp_path):
    m = AstroPhot_Model(name=None, filename=GOLDEN)
    assert type(m).__name__ == "Sersic_Galaxy" and m.name == "test sersic"
    assert m["center"].value.tolist() == [20.25, 19.5]
    assert float(m["PA"].value) == pytest.approx(np.pi / 3, rel=1e-15) and m["PA"].cyclic
    assert m["n"].locked and m["n"].limits == (0.36, 8.0)
    assert m["Re"].limits == (0.0, None)
    assert m.psf_mode == "full" and m.integrate_mode == "none" and m.sampling_tolerance == 1e-3
    assert m.window.nx == 50 and m.window.ny == 40
    assert torch.allclose(m.window.pixelscale, torch.tensor([[0.8, 0.0], [0.0, 0.8]], dtype=torch.float64))
    assert m.psf.data.shape == (5, 5) and float(m.psf.data.sum()) == pytest.approx(0.75)
    # identities are kept: a group that links this model keeps its columns
    ref = yaml.safe_load(open(GOLDEN))
    assert m["Re"].identity == ref["parameters"]["nodes"][4]["identity"]
    out = str(tmp_path / "again.yaml")
    m.save(out)
    back = yaml.safe_load(open(out))
    ref.pop("target_identity")
    back.pop("target_identity")
    assert back == ref


def test_safe_loader_builds_no_python_objects(tmp_path):
    fn = tmp_path / "evil.yaml"
    fn.write_text("name: x\nmodel_type: !!python/object/apply:os.getcwd []\n")
    with pytest.raises(yaml.YAMLError):
        AstroPhot_Model.load(str(fn))
    with pytest.raises(yaml.YAMLError):
        AstroPhot_Model.load(io.StringIO(fn.read_text()))


def test_unknown_format_and_model_type(tmp_path):
    m = sersic("s", target())
    with pytest.raises(ValueError):
        m.save(str(tmp_path / "m.txt"))
    st = m.get_state()
    st["model_type"] = "no such model"
    with pytest.raises(ap.errors.UnrecognizedModel):
        AstroPhot_Model(name="x", filename=st)
    assert json.loads(json.dumps(m.get_state()))["model_type"] == "sersic galaxy model"


def test_group_with_psf_model_state(tmp_path):
    """The reference's group save/load test (tests/test_group_models.py:112-165):
    a Moffat PSF model (n, Rd free) as the first model's PSF gives 16
    parameters, its node leading that model's branch; all of it survives
    save and load."""
    tar = target(51, 51, 1.0)
    psf = AstroPhot_Model(name="psf model 1", model_type="moffat psf model",
                          target=ap.image.PSF_Image(data=np.zeros((11, 11)), pixelscale=1.0),
                          parameters={"n": 2.0, "Rd": 3.0})
    mod1 = sersic("base model 1", tar, psf=psf, psf_mode="full")
    mod2 = sersic("base model 2", tar)
    smod = AstroPhot_Model(name="group model", model_type="group model", models=[mod1, mod2], target=tar)
    assert smod.parameters.size == 16
    assert list(mod1.parameters.vector_names())[:2] == ["n", "Rd"]
    fn = str(tmp_path / "g.yaml")
    smod.save(fn)
    new = AstroPhot_Model(name="group model", filename=fn)
    assert len(new.models) == 2 and new.parameters.size == 16
    assert torch.equal(new.parameters.vector_values(), smod.parameters.vector_values())
    p1 = new.models["base model 1"].psf
    assert type(p1).__name__ == "Moffat_PSF" and p1.window.get_state() == psf.window.get_state()
    assert p1.parameters is new.models["base model 1"].parameters["psf model 1"]
    assert p1["I0"].locked and p1["center"].locked
    assert [s for s, _ in new.models["base model 1"]._free_slots()] == \
        [s for s, _ in mod1._free_slots()]


def test_psf_model_must_be_centred_and_unconvolved():
    img = ap.image.PSF_Image(data=np.zeros((9, 9)), pixelscale=0.5)
    m = AstroPhot_Model(name="p", model_type="moffat psf model", target=img, parameters={"n": 2.0, "Rd": 1.0})
    assert m.psf_mode == "none" and m.sampling_mode == "simpsons" and m.sampling_tolerance == 1e-3
    assert m.parameters.size == 2 and m.window.nx == 9
    with pytest.raises(ap.errors.SpecificationConflict):
        m.psf_mode = "full"
    with pytest.raises(ap.errors.InvalidTarget):
        AstroPhot_Model(name="q", model_type="moffat psf model", target=target(), parameters={"n": 2.0, "Rd": 1.0})
    with pytest.raises(ap.errors.SpecificationConflict):
        AstroPhot_Model(name="r", model_type="moffat psf model", target=img,
                        window=ap.image.Window(pixel_shape=(9, 9), pixelscale=0.5, origin=(0.0, 0.0)),
                        parameters={"n": 2.0, "Rd": 1.0})
