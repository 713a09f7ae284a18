"""
Event log kept by the integration and phi-manipulation calls.

dadi appends one record per call to ``dadi.Demes.cache``
(dadi/Integration.py:216, 257, 322, 386, 469, 553, 726, 936;
dadi/PhiManip.py:33, 230, 388) so that a model function's demography can be
exported afterwards.  The export itself (demes graphs) is not part of this
package; the records are kept with the same fields so callers that inspect
the cache keep working.
"""
import numpy as np

cache = []


class Event:
    def __init__(self, duration=0, deme_ids=None):
        self.duration = duration
        self.end_time = None
        self.deme_ids = tuple(deme_ids) if deme_ids is not None else None


class Initiation(Event):
    def __init__(self, nu, deme_ids=None):
        super().__init__(np.inf, deme_ids=deme_ids)
        self.start_sizes = [nu]
        self.end_sizes = None


class IntegrationConst(Event):
    def __init__(self, duration, start_sizes, mig=(), deme_ids=None):
        super().__init__(duration, deme_ids=deme_ids)
        self.start_sizes = list(start_sizes)
        self.end_sizes = None
        self.mig = list(mig)


class IntegrationNonConst(Event):
    def __init__(self, history, deme_ids=None):
        super().__init__(duration=history[-1][0] - history[0][0], deme_ids=deme_ids)
        self.start_sizes = history[0][1]
        self.end_sizes = history[-1][1]
        self.mig = history[0][2]
        self.history = history


class Split(Event):
    def __init__(self, proportions, deme_ids=None):
        super().__init__(deme_ids=deme_ids)
        self.proportions = proportions


class Remove(Event):
    def __init__(self, removed):
        super().__init__()
        self.removed = removed
