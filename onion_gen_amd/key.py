"""Key derivation and hit reconstruction (src/key.rs, src/run.rs:211), the
batched hit sink and the key-file writer (src/output.rs)."""
import ctypes
from dataclasses import dataclass

from ._lib import OgError, err_buf, lib

HIT_OK, HIT_CLAMPING, HIT_MISMATCH = 0, 1, 2


class _KeyRecord(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("public_key", ctypes.c_uint8 * 32), ("expanded_secret", ctypes.c_uint8 * 64),
                ("hostname", ctypes.c_char * 63)]


@dataclass(frozen=True)
class KeyRecord:
    offset: int
    status: int
    public_key: bytes
    expanded_secret: bytes
    hostname: str


def secret_scalar(seed):
    out = ctypes.create_string_buffer(32)
    lib().og_key_secret_scalar(bytes(seed), out)
    return out.raw


def expanded_secret(seed):
    out = ctypes.create_string_buffer(64)
    lib().og_key_expanded_secret(bytes(seed), out)
    return out.raw


def start_point(seed):
    """s0*B of a seed as 128 bytes: X, Y, Z, T."""
    out = ctypes.create_string_buffer(128)
    lib().og_key_start_point(bytes(seed), out)
    return out.raw


def public_from_scalar(scalar):
    out = ctypes.create_string_buffer(32)
    lib().og_key_public_from_scalar(bytes(scalar), out)
    return out.raw


def expanded_secret_at_offset(seed, offset):
    """The 64-byte secret of the candidate at ``offset``, or None when the
    offset broke clamping (src/key.rs:167-177)."""
    out = ctypes.create_string_buffer(64)
    ok = lib().og_key_expanded_secret_at_offset(bytes(seed), offset, out)
    return out.raw if ok else None


def address_bytes(pk):
    out = ctypes.create_string_buffer(35)
    lib().og_key_address_bytes(bytes(pk), out)
    return out.raw


def address(pk):
    out = ctypes.create_string_buffer(57)
    lib().og_key_address(bytes(pk), out)
    return out.value.decode(errors="replace")


def block_seed(root, index):
    out = ctypes.create_string_buffer(32)
    lib().og_run_block_seed(bytes(root), index, out)
    return out.raw


def finish_hit(seed, offset):
    """(pk with sign, address) of the candidate at ``offset``, re-derived from
    its secret scalar independently of the chain that found it
    (src/run.rs:1089-1104); None when clamping broke."""
    secret = expanded_secret_at_offset(seed, offset)
    if secret is None:
        return None
    pk = public_from_scalar(secret[:32])
    return pk, address(pk)


def finish_hits(seed, survivors, device=0):
    """The batched hit sink (og_hip_finish_hits): for each survivor (objects
    with ``offset`` and ``packed``) its secret at the offset, its public key
    re-derived on the device from the scalar alone and compared with the
    survivor, and its hostname.  Returns KeyRecords in the same order."""
    from .gpu import _Survivor
    n = len(survivors)
    if n == 0:
        return []
    src = (_Survivor * n)()
    for i, s in enumerate(survivors):
        src[i].offset = s.offset
        src[i].packed[:] = list(s.packed)
    out = (_KeyRecord * n)()
    e = err_buf()
    if lib().og_hip_finish_hits(device, bytes(seed), src, n, out, e, len(e)) != 0:
        raise OgError(e.value.decode(errors="replace"))
    return [KeyRecord(r.offset, r.status, bytes(r.public_key), bytes(r.expanded_secret),
                      r.hostname.decode()) for r in out]


def write_key(directory, public_key, expanded_secret):
    """Writes tor's key directory for one key (og_write_key).  Returns
    (created, path); an existing directory is left as it is."""
    path = ctypes.create_string_buffer(4096)
    rc = lib().og_write_key(str(directory).encode(), bytes(public_key), bytes(expanded_secret),
                            path, len(path))
    if rc < 0:
        raise OSError(path.value.decode(errors="replace"))
    return rc == 0, path.value.decode(errors="replace")
