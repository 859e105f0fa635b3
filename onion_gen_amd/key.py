"""Key derivation and hit reconstruction (src/key.rs, src/run.rs:211)."""
import ctypes

from ._lib import lib


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
    return out.value.decode()


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
