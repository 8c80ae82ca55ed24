"""imageproc_amd: imageproc's sliding-window correlation path on MI355X (gfx950).

Host-side mirror of imageproc 0.27.0's ``filter`` and ``template_matching``
modules over the C ABI of ``libimageproc_hip.so`` (include/imageproc_hip.h):

    from imageproc_amd import filter, template_matching, Kernel
    blurred = filter.gaussian_blur_f32(gray, 2.0)
    scores = template_matching.match_template(
        gray, tmpl, template_matching.MatchTemplateMethod.CrossCorrelationNormalized)

``imageproc_amd.device`` holds the device-resident (torch tensor) entry points.
"""
from . import filter, kernel, template_matching
from ._lib import DeviceError, Panic, device_info
from .kernel import Kernel
from .template_matching import (Extremes, MatchTemplateMethod, find_extremes, match_template,
                                match_template_with_mask)

__all__ = ["filter", "kernel", "template_matching", "Kernel", "Panic", "DeviceError",
           "MatchTemplateMethod", "match_template", "find_extremes", "Extremes", "device_info"]
