"""``numba.types`` stand-in (names only)."""


class StringLiteral:
    pass


class BaseTuple:
    @staticmethod
    def from_types(args):
        return tuple(args)


def void(*a):
    return None


float64 = float
int64 = int
boolean = bool
