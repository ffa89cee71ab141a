"""Step-chain runner: builds the registered step closures for a named chain
of the package's pipeline_config.yml and applies them to a RubixData, the
part of RubixPipeline.run (rubix/core/pipeline.py:93-201) that covers this
path.  Chains: "calc_ifu" (rotate_galaxy -> filter_particles -> the
per-step chain -> convolve_lsf -> apply_noise, the reference's chain),
"calc_ifu_fused" (the same with the fused datacube step), "calc_dusty_ifu"
(calc_ifu with calculate_extinction after the resampling), "calc_dusty_ifu_fused"
(the same with one fused dusty datacube step), and
"calc_ifu_aligned" / "calc_ifu_fused_aligned" from spaxel_assignment to
convolve_lsf for particles that are already aligned and filtered."""
from __future__ import annotations

from ..config import pipeline_config
from ..pipeline import LinearTransformerPipeline
from .data import RubixData, get_reshape_data, get_rubix_data
from .ifu import (get_calculate_datacube, get_calculate_datacube_fused, get_calculate_spectra,
                  get_doppler_shift_and_resampling, get_scale_spectrum_by_mass)
from .dust import get_calculate_dusty_datacube_fused, get_extinction
from .lsf import get_convolve_lsf
from .noise import get_apply_noise
from .psf import get_convolve_psf
from .rotation import get_galaxy_rotation
from .telescope import get_filter_particles, get_spaxel_assignment

FACTORIES = {
    "rotate_galaxy": get_galaxy_rotation,
    "filter_particles": get_filter_particles,
    "spaxel_assignment": get_spaxel_assignment,
    "reshape_data": get_reshape_data,
    "calculate_spectra": get_calculate_spectra,
    "scale_spectrum_by_mass": get_scale_spectrum_by_mass,
    "doppler_shift_and_resampling": get_doppler_shift_and_resampling,
    "calculate_extinction": get_extinction,
    "calculate_datacube": get_calculate_datacube,
    "calculate_datacube_fused": get_calculate_datacube_fused,
    "calculate_dusty_datacube_fused": get_calculate_dusty_datacube_fused,
    "convolve_psf": get_convolve_psf,
    "convolve_lsf": get_convolve_lsf,
    "apply_noise": get_apply_noise,
}


class RubixPipeline:
    def __init__(self, config: dict, pipeline_name: str = None):
        self.config = config
        name = pipeline_name or config.get("pipeline", {}).get("name", "calc_ifu")
        if name not in pipeline_config:
            raise ValueError(f"Pipeline {name} not found in the pipeline configuration")
        self.pipeline_name = name
        chain = pipeline_config[name]
        fns = []
        for node in chain["Transformers"].values():
            step = node["name"]
            if step not in FACTORIES:
                raise ValueError(f"Step {step} is not implemented on the GPU path")
            fns.append(FACTORIES[step](config))
        self.pipeline = LinearTransformerPipeline(chain, fns)

    def run(self, rubixdata: RubixData = None) -> RubixData:
        """Runs the chain on rubixdata, or, as the reference's
        RubixPipeline.run does, on the galaxy in
        config["output_path"]/rubix_galaxy.h5 (rubix/core/pipeline.py:52-78,
        138-175; get_rubix_data)."""
        if rubixdata is None:
            rubixdata = get_rubix_data(self.config)
        return self.pipeline.apply(rubixdata)

    def gradient(self, rubixdata: RubixData, loss_fn):
        """Gradients of loss_fn(datacube) w.r.t. every star field that
        requires grad (the reference raises NotImplementedError here,
        rubix/core/pipeline.py:203-208)."""
        out = self.run(rubixdata)
        loss = loss_fn(out.stars.datacube)
        loss.backward()
        return loss
