//! Declarations of the C ABI this crate binds: the subset of `include/zfp.h`
//! (zfp-rs-ffi's header, zfp-rs-ffi/include/zfp.h:163-365) and
//! `include/zfp_hip.h` that the safe API in `lib.rs` calls.  The struct
//! layouts are the header's.
#![allow(non_camel_case_types)]

use core::ffi::c_void;

pub type zfp_type = u32;
pub const ZFP_TYPE_INT32: zfp_type = 1;
pub const ZFP_TYPE_INT64: zfp_type = 2;
pub const ZFP_TYPE_FLOAT: zfp_type = 3;
pub const ZFP_TYPE_DOUBLE: zfp_type = 4;

pub const ZFP_ROUND_FIRST: i32 = -1;
pub const ZFP_ROUND_NEVER: i32 = 0;
pub const ZFP_ROUND_LAST: i32 = 1;

/// `bitstream` is opaque.
#[repr(C)]
pub struct bitstream {
    _private: [u8; 0],
}

#[repr(C)]
pub struct zfp_execution {
    pub policy: u32,
    pub params: *mut c_void,
}

#[repr(C)]
pub struct zfp_stream {
    pub minbits: u32,
    pub maxbits: u32,
    pub maxprec: u32,
    pub minexp: i32,
    pub stream: *mut bitstream,
    pub exec: zfp_execution,
}

#[repr(C)]
pub struct zfp_field {
    pub type_: zfp_type,
    pub nx: usize,
    pub ny: usize,
    pub nz: usize,
    pub nw: usize,
    pub sx: isize,
    pub sy: isize,
    pub sz: isize,
    pub sw: isize,
    pub data: *mut c_void,
}

extern "C" {
    pub fn stream_open(buffer: *mut c_void, bytes: usize) -> *mut bitstream;
    pub fn stream_close(s: *mut bitstream);
    pub fn stream_capacity(s: *const bitstream) -> usize;
    pub fn stream_rseek(s: *mut bitstream, offset: u64);
    pub fn stream_wseek(s: *mut bitstream, offset: u64);
    pub fn stream_rtell(s: *const bitstream) -> u64;
    pub fn stream_wtell(s: *const bitstream) -> u64;

    pub fn zfp_stream_open(s: *mut bitstream) -> *mut zfp_stream;
    pub fn zfp_stream_close(z: *mut zfp_stream);
    pub fn zfp_stream_set_bit_stream(z: *mut zfp_stream, s: *mut bitstream);
    pub fn zfp_stream_set_rate(z: *mut zfp_stream, rate: f64, t: zfp_type, dims: u32, align: i32) -> f64;
    pub fn zfp_stream_set_precision(z: *mut zfp_stream, precision: u32) -> u32;
    pub fn zfp_stream_set_accuracy(z: *mut zfp_stream, tolerance: f64) -> f64;
    pub fn zfp_stream_set_reversible(z: *mut zfp_stream);
    pub fn zfp_stream_set_params(z: *mut zfp_stream, minbits: u32, maxbits: u32, maxprec: u32, minexp: i32) -> i32;
    pub fn zfp_stream_maximum_size(z: *const zfp_stream, f: *const zfp_field) -> usize;
    pub fn zfp_compress(z: *mut zfp_stream, f: *const zfp_field) -> usize;
    pub fn zfp_decompress(z: *mut zfp_stream, f: *mut zfp_field) -> usize;

    // include/zfp_hip.h
    pub fn zfp_stream_set_rounding(z: *mut zfp_stream, rounding: i32, tight_error: i32) -> i32;
    pub fn zfp_stream_set_hip_stream(z: *mut zfp_stream, hip_stream: *mut c_void) -> i32;
    pub fn zfp_stream_index_export(s: *const bitstream, lengths: *mut u32, capacity: usize, start: *mut u64) -> usize;
    pub fn zfp_stream_index_import(s: *mut bitstream, lengths: *const u32, count: usize, start: u64) -> i32;
}
