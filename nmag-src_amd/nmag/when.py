"""Event schedules for the `save` and `do` lists of relax / hysteresis.

Follows the semantics of interface/nsim/when.py:
  at(id, value)                      :11-25
  every(id, delta, first, last)      :35-63, :208-236
  a | b, a & b                       :69-120, :176-199
  When.next_time with a tolerance    :163-174
  never                              :238

A schedule is asked two things about a clock (a mapping with 'step',
'stage_step', 'stage', 'time', 'stage_time', 'real_time' and the flags
'convergence', 'stage_end'):

* match_time(clock): does the event fall on the clock's current values;
* next_time(identifier, clock): the next value of that counter at which the
  event falls.  The answer is a number, or a bool when the schedule does not
  constrain that counter: True means "any value" and False "never".

Time values may be given as SI objects; the clock holds SI seconds as floats.
"""
from .units import SI


def _plain(v):
    return v.value if isinstance(v, SI) else v


def _is_whole(x):
    # the reference's float_is_integer (when.py:4-5): x within half an ulp of 1.0
    # of an integer
    return 1.0 + (x - round(x)) * 0.5 == 1.0


class When:
    """Base of the schedule nodes; combine with | and &."""

    def match_time(self, clock):
        raise NotImplementedError

    def next_time(self, identifier, clock, tols=None):
        nt = self._next(identifier, clock)
        if tols and identifier in tols and not isinstance(nt, bool):
            tol = tols[identifier]
            now = clock[identifier]
            if tol > 0.0 and abs(nt - now) < tol:
                # an event at (about) the current value is not the next one
                shifted = dict(clock)
                shifted[identifier] = nt + tol
                return self._next(identifier, shifted)
        return nt

    def _next(self, identifier, clock):
        raise NotImplementedError

    def __or__(self, other):
        return _Or(self, other)

    def __and__(self, other):
        return _And(self, other)

    def __str__(self):
        return repr(self)


class _At(When):
    def __init__(self, identifier, value):
        self.identifier, self.value = identifier, _plain(value)

    def __repr__(self):
        return "at(%s, %s)" % (self.identifier, self.value)

    def match_time(self, clock):
        return clock[self.identifier] == self.value

    def _next(self, identifier, clock):
        if identifier != self.identifier:
            return True
        now = clock[identifier]
        if isinstance(now, bool):
            return now == self.value
        return self.value if now < self.value else False


class _Every(When):
    def __init__(self, identifier, delta, first, last):
        self.identifier = identifier
        self.delta, self.first, self.last = _plain(delta), _plain(first), _plain(last)

    def __repr__(self):
        opts = "" if self.first == 0 else ", first=%s" % self.first
        opts += "" if self.last is None else ", last=%s" % self.last
        return "every('%s', %s%s)" % (self.identifier, self.delta, opts)

    def match_time(self, clock):
        t = clock[self.identifier]
        if self.first is not None and t < self.first:
            return False
        if self.last is not None and t > self.last:
            return False
        if self.delta is None:
            return True
        return _is_whole((t - self.first) / self.delta)

    def _next(self, identifier, clock):
        if identifier != self.identifier:
            return self.match_time(clock)
        t = clock[identifier]
        if self.last is not None and t >= self.last:
            return False
        if self.first is not None and t < self.first and self.delta is not None:
            return self.first
        if self.delta is None:
            return True
        k = (t - self.first) / self.delta
        n = int(round(k)) + 1 if _is_whole(k) else int(k) + 1
        nxt = self.first + self.delta * n
        if self.last is not None and nxt > self.last:
            return False
        return nxt


class _Or(When):
    def __init__(self, a, b):
        self.a, self.b = a, b

    def __repr__(self):
        return "%r | %r" % (self.a, self.b)

    def match_time(self, clock):
        return self.a.match_time(clock) or self.b.match_time(clock)

    def _next(self, identifier, clock):
        na, nb = self.a._next(identifier, clock), self.b._next(identifier, clock)
        if isinstance(na, bool) and isinstance(nb, bool):
            return na or nb
        if isinstance(na, bool):
            return nb
        if isinstance(nb, bool):
            return na
        return min(na, nb)


class _And(When):
    """Both events at once.  Found by stepping both schedules forward until
    they agree, as the reference does (when.py:92-120): only meaningful for
    schedules over the same counter, and bounded here so that an empty
    intersection ends instead of looping."""

    max_probes = 100000

    def __init__(self, a, b):
        self.a, self.b = a, b

    def __repr__(self):
        return "%r & %r" % (self.a, self.b)

    def match_time(self, clock):
        return self.a.match_time(clock) and self.b.match_time(clock)

    def _next(self, identifier, clock):
        probe = dict(clock)
        for _ in range(self.max_probes):
            na, nb = self.a._next(identifier, probe), self.b._next(identifier, probe)
            if na is False or nb is False:
                return False
            if na is True:
                return nb
            if nb is True:
                return na
            # jump to the later candidate; done when the other schedule
            # falls there too, else search on from it
            later, other = (na, self.b) if na > nb else (nb, self.a)
            probe[identifier] = later
            if other.match_time(probe):
                return later
        return False


class _Never(When):
    def __repr__(self):
        return "never"

    def match_time(self, clock):
        return False

    def _next(self, identifier, clock):
        return False


never = _Never()


def at(identifier, value=True):
    """The event at one value of a counter or flag: at('convergence'),
    at('stage_end'), at('step', 10), at('time', SI(1e-10, 's'))."""
    return _At(identifier, value)


def every(arg1, arg2=None, first=0, last=None):
    """The periodic event every('step', 10), every('time', SI(1e-12, 's')),
    every('step', 5, first=10, last=100); every(10, 'step') is accepted too
    (interface/nsim/when.py:208-236)."""
    if isinstance(arg1, str):
        identifier, delta = arg1, arg2
    else:
        identifier, delta = arg2, arg1
    if not isinstance(identifier, str):
        raise ValueError("every() needs a counter name, e.g. every('step', 10)")
    if last is not None and _plain(last) <= _plain(first):
        raise ValueError("every(): last (%s) must be greater than first (%s)" % (last, first))
    if delta is not None and _plain(delta) <= 0:
        raise ValueError("every(): the interval must be positive")
    return _Every(identifier, delta, first, last)


def next_deltas(event, clock, suggest=None, tols=None):
    """(steps, simulation time, real time) from the clock to the event's next
    occurrence, each None when the event does not bound it, and each no later
    than `suggest` (interface/nmag/hysteresis.py:136-160)."""

    def delta(name):
        n = event.next_time(name, clock, tols=tols)
        return None if isinstance(n, bool) else n - clock[name]

    def least(*vals):
        vals = [v for v in vals if v is not None]
        return min(vals) if vals else None

    s_step, s_time, s_real = suggest if suggest is not None else (None, None, None)
    return (least(delta("step"), delta("stage_step"), s_step),
            least(delta("time"), delta("stage_time"), s_time),
            least(delta("real_time"), s_real))


def next_occurrence(event, clock, tols=None):
    """The event's next (stage step, stage time, real time), the key the
    relax/hysteresis loop compares before and after an advance to tell that
    an event was passed (interface/nmag/hysteresis.py:162-200)."""

    def nxt(name, origin=None):
        n = event.next_time(name, clock, tols=tols)
        if isinstance(n, bool):
            return None
        return n - clock[origin] if origin is not None else n

    def least(*vals):
        vals = [v for v in vals if v is not None]
        return min(vals) if vals else None

    return (least(nxt("step", "zero_stage_step"), nxt("stage_step")),
            least(nxt("time", "zero_stage_time"), nxt("stage_time")),
            nxt("real_time"))
