"""PIV errors (same contract as src/fluidimage/calcul/errors.py:8-15)."""


class PIVError(Exception):
    """No peak (or no usable sub-pixel fit) for one window.

    Carries ``explanation`` (the reference's message) and, when raised from
    the peak search, ``results = (dx, dy, correl_max)``.
    """

    def __init__(self, *args, **kargs):
        self.explanation = "General no peak error"
        for k, v in kargs.items():
            self.__dict__[k] = v
        super().__init__(*args)
