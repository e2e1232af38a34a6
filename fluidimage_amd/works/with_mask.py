"""Image masking by ``params.mask.strcrop`` (src/fluidimage/works/with_mask.py:15-47,
src/fluidimage/data_objects/piv.py:71-75)."""


def slices_from_strcrop(strcrop):
    return tuple(
        slice(*(int(i) if i else None for i in part.strip().split(":")))
        for part in strcrop.split(",")
    )


class BaseWorkWithMask:
    params = None

    def _shift(self, xs, ys, sign):
        strcrop = self.params.mask.strcrop
        if strcrop is not None:
            slices = slices_from_strcrop(strcrop)
            if slices[1].start is not None:
                xs = xs + sign * slices[1].start
            if slices[0].start is not None:
                ys = ys + sign * slices[0].start
        return xs, ys

    def _xymasked_from_xyoriginalimage(self, xs, ys):
        return self._shift(xs, ys, -1)

    def _xyoriginalimage_from_xymasked(self, xs, ys):
        return self._shift(xs, ys, +1)

    @classmethod
    def _complete_params_with_default_mask(cls, params):
        params._set_child("mask", attribs={"strcrop": None})
