"""Frames whose encoder-side decisions come from the reference encoder: the
quant field from InitialQuantField (pik/adaptive_quantization.cc:1047-1060)
through Quantizer::SetQuantField's rule, the colour correlation from
FindBestColorCorrelationMap (pik/color_correlation.h:101), and the
quantized AC / DC from the reference's own quantization
(InitializeFrameEncCache + ComputeCoefficients, as RoundtripImage runs them)
of the Gaborish-sharpened XYB image. Only the AC strategy map stays the
generator's (the reference's strategy search needs colour management that
is not buildable here). Test helper: it calls the reference harness."""
import numpy as np

from pik_amd import synth


def reference_side_frame(w, h, seed, distance=1.0, predict=True,
                         num_rects=40):
    from oracle import ref
    fr, opsin, orig, _ = synth.make_search_inputs(w, h, seed=seed,
                                                  num_rects=num_rects,
                                                  target=distance)
    fr.predict_lf = fr.predict_hf = int(predict)
    init = ref.initial_quant_field(orig, distance)
    gs, qdc, field = synth.set_quant_field(
        init, ref.initial_quant_dc(distance))
    # a multi-block transform takes its blocks' maximum (AdjustQuantField)
    raw = fr.ac_strategy >> 4
    for code, s in ((4, 2), (5, 4)):
        ys, xs = np.nonzero((raw == code) & ((fr.ac_strategy & 15) == 0))
        for y0, x0 in zip(ys, xs):
            field[y0:y0 + s, x0:x0 + s] = field[y0:y0 + s, x0:x0 + s].max()
    fr.global_scale, fr.quant_dc, fr.quant_field = gs, qdc, field
    pad = np.ascontiguousarray(opsin)
    ytox, ytob, xdc, bdc = ref.color_correlation(pad)
    fr.ytox_map, fr.ytob_map, fr.ytox_dc, fr.ytob_dc = ytox, ytob, xdc, bdc
    ac, dc = ref.quantize(fr, opsin)
    fr.ac, fr.dc = ac, dc
    return fr.contiguous()
