//! zfp-rs's compress/decompress path on an MI355X.
//!
//! A safe binding to `libzfp_hip.so`, the library that exports zfp-rs-ffi's C
//! ABI with every block coded by HIP kernels for gfx950.  It keeps zfp-rs's
//! names and argument meaning for the path it covers:
//!
//! - [`ZfpRounding`] and [`ZfpConfig`] mirror `src/config.rs:115-339`
//!   (the mode constructors, `with_rounding`, the expert parameters);
//! - [`ZfpField`] / [`ZfpFieldMut`] mirror `src/field.rs:285-460` for
//!   contiguous 1-D to 4-D slices (`new`) and strided views (`new_strided`);
//! - [`HipBitStream::compress`] / [`HipBitStream::decompress`] mirror
//!   `ZfpBitStream::compress` / `decompress` (`src/bitstream/ops.rs:98-234`):
//!   they return the stream size (compress) or read position (decompress)
//!   in bytes, word aligned, and the same error variants.
//!
//! Streams and fields may live in host memory or in device memory allocated
//! with `hipMalloc` (a pointer the HIP runtime knows is coded in place).
//!
//! This crate is source only in the repository that builds the library: that
//! environment has no Rust toolchain, so it has not been compiled there.
//! `build.rs` links `libzfp_hip.so` from `ZFP_HIP_DIR`.

mod ffi;

use core::ffi::c_void;
use core::marker::PhantomData;

/// Coefficient rounding (zfp-rs `src/config.rs:115-149`); a per-call
/// setting that is not part of the stream's mode word.
#[derive(Debug, Clone, Copy, PartialEq, Eq, Hash, Default)]
pub enum ZfpRounding {
    #[default]
    Never,
    First { tight_error: bool },
    Last { tight_error: bool },
}

impl ZfpRounding {
    pub fn tight_error(self) -> bool {
        match self {
            Self::Never => false,
            Self::First { tight_error } | Self::Last { tight_error } => tight_error,
        }
    }

    fn c_mode(self) -> i32 {
        match self {
            Self::Never => ffi::ZFP_ROUND_NEVER,
            Self::First { .. } => ffi::ZFP_ROUND_FIRST,
            Self::Last { .. } => ffi::ZFP_ROUND_LAST,
        }
    }
}

/// Scalar types of a field (zfp-rs `ZfpScalarType`).
#[derive(Debug, Clone, Copy, PartialEq, Eq, Hash)]
pub enum ZfpScalarType {
    Int32,
    Int64,
    Float,
    Double,
}

impl ZfpScalarType {
    fn c_type(self) -> ffi::zfp_type {
        match self {
            Self::Int32 => ffi::ZFP_TYPE_INT32,
            Self::Int64 => ffi::ZFP_TYPE_INT64,
            Self::Float => ffi::ZFP_TYPE_FLOAT,
            Self::Double => ffi::ZFP_TYPE_DOUBLE,
        }
    }
}

/// Element types a field may hold (zfp-rs `ZfpScalar`).
pub trait ZfpScalar: Copy + sealed::Sealed {
    const TYPE: ZfpScalarType;
}
impl ZfpScalar for i32 {
    const TYPE: ZfpScalarType = ZfpScalarType::Int32;
}
impl ZfpScalar for i64 {
    const TYPE: ZfpScalarType = ZfpScalarType::Int64;
}
impl ZfpScalar for f32 {
    const TYPE: ZfpScalarType = ZfpScalarType::Float;
}
impl ZfpScalar for f64 {
    const TYPE: ZfpScalarType = ZfpScalarType::Double;
}
mod sealed {
    pub trait Sealed {}
    impl Sealed for i32 {}
    impl Sealed for i64 {}
    impl Sealed for f32 {}
    impl Sealed for f64 {}
}

/// Expert parameters plus rounding (zfp-rs `ZfpConfig`,
/// `src/config.rs:163-339`).  The mode constructors take their arithmetic
/// from the library's C entry points, which restate zfp-rs's.
#[derive(Debug, Clone, Copy, PartialEq, Eq, Hash)]
pub struct ZfpConfig {
    min_bits: u32,
    max_bits: u32,
    max_prec: u32,
    min_exp: i32,
    rounding: ZfpRounding,
}

/// Runs `f` on a scratch zfp_stream and reads back its expert parameters.
fn params_from(f: impl FnOnce(*mut ffi::zfp_stream)) -> ZfpConfig {
    unsafe {
        let z = ffi::zfp_stream_open(core::ptr::null_mut());
        assert!(!z.is_null(), "zfp_stream_open");
        f(z);
        let c = ZfpConfig {
            min_bits: (*z).minbits,
            max_bits: (*z).maxbits,
            max_prec: (*z).maxprec,
            min_exp: (*z).minexp,
            rounding: ZfpRounding::Never,
        };
        ffi::zfp_stream_close(z);
        c
    }
}

impl ZfpConfig {
    /// Fixed rate in bits per value for `dims`-dimensional fields of `ty`
    /// (`config.rs:266-306`); `word_aligned` rounds blocks up to whole words.
    pub fn fixed_rate(rate: f64, ty: ZfpScalarType, dims: u32, word_aligned: bool) -> Self {
        params_from(|z| unsafe {
            ffi::zfp_stream_set_rate(z, rate, ty.c_type(), dims, word_aligned as i32);
        })
    }

    pub fn fixed_precision(precision: u32) -> Self {
        params_from(|z| unsafe {
            ffi::zfp_stream_set_precision(z, precision);
        })
    }

    pub fn fixed_accuracy(tolerance: f64) -> Self {
        params_from(|z| unsafe {
            ffi::zfp_stream_set_accuracy(z, tolerance);
        })
    }

    pub fn reversible() -> Self {
        params_from(|z| unsafe { ffi::zfp_stream_set_reversible(z) })
    }

    /// Expert parameters; `None` when they are out of range
    /// (`valid_params`, `config.rs:103-106`).
    pub fn expert(min_bits: u32, max_bits: u32, max_prec: u32, min_exp: i32) -> Option<Self> {
        (min_bits <= max_bits && (1..=64).contains(&max_prec)).then_some(Self {
            min_bits,
            max_bits,
            max_prec,
            min_exp,
            rounding: ZfpRounding::Never,
        })
    }

    pub fn with_rounding(mut self, rounding: ZfpRounding) -> Self {
        self.rounding = rounding;
        self
    }

    pub fn min_bits(&self) -> u32 {
        self.min_bits
    }
    pub fn max_bits(&self) -> u32 {
        self.max_bits
    }
    pub fn max_prec(&self) -> u32 {
        self.max_prec
    }
    pub fn min_exp(&self) -> i32 {
        self.min_exp
    }
    pub fn rounding(&self) -> ZfpRounding {
        self.rounding
    }

    /// Bytes that always hold the stream of `field` (`config.rs:642-686`).
    pub fn maximum_size(&self, field: &ZfpField<'_>) -> usize {
        let z = Stream::scratch(self);
        let n = unsafe { ffi::zfp_stream_maximum_size(z.0, &field.raw) };
        n
    }
}

/// Errors of `compress` (zfp-rs `ZfpCompressionError`, `src/types.rs:195-216`,
/// for the variants this path can give).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum ZfpCompressionError {
    /// The field or stream could not be coded: an invalid shape, or the GPU
    /// call failed (nothing was written to the stream).
    Field,
    /// The output did not fit: writes past the capacity were dropped.
    BufferTooSmall { required: usize, capacity: usize },
}

/// Errors of `decompress` (zfp-rs `ZfpDecompressionError`).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum ZfpDecompressionError {
    Field,
    /// Decoding read a word past the end of the buffer.
    Truncated,
}

/// A read-only field over a slice, x fastest (zfp-rs `ZfpField`).
pub struct ZfpField<'a> {
    raw: ffi::zfp_field,
    _data: PhantomData<&'a [u8]>,
}

/// A writable field (zfp-rs `ZfpFieldMut`).
pub struct ZfpFieldMut<'a> {
    raw: ffi::zfp_field,
    _data: PhantomData<&'a mut [u8]>,
}

/// Dimensions (x first) and strides in elements; zero strides mean
/// contiguous.  `None` when a size is zero, there are more than four, or
/// `len` elements do not cover the view.
fn describe<T: ZfpScalar>(
    ptr: *mut c_void,
    len: usize,
    dims: &[usize],
    strides: Option<&[isize]>,
) -> Option<ffi::zfp_field> {
    if dims.is_empty() || dims.len() > 4 || dims.contains(&0) {
        return None;
    }
    let mut n = [0usize; 4];
    let mut s = [0isize; 4];
    n[..dims.len()].copy_from_slice(dims);
    let mut span_lo: isize = 0;
    let mut span_hi: isize = 0;
    let mut natural: isize = 1;
    for a in 0..dims.len() {
        let st = match strides {
            Some(v) => *v.get(a)?,
            None => natural,
        };
        s[a] = if strides.is_some() { st } else { 0 };
        let reach = st.checked_mul(isize::try_from(dims[a] - 1).ok()?)?;
        if reach < 0 {
            span_lo = span_lo.checked_add(reach)?;
        } else {
            span_hi = span_hi.checked_add(reach)?;
        }
        natural = natural.checked_mul(isize::try_from(dims[a]).ok()?)?;
    }
    // a view must stay inside the slice; negative strides start past its
    // beginning (the origin is the element of index 0)
    if span_lo != 0 || usize::try_from(span_hi).ok()? >= len {
        return None;
    }
    Some(ffi::zfp_field {
        type_: T::TYPE.c_type(),
        nx: n[0],
        ny: n[1],
        nz: n[2],
        nw: n[3],
        sx: s[0],
        sy: s[1],
        sz: s[2],
        sw: s[3],
        data: ptr,
    })
}

impl<'a> ZfpField<'a> {
    /// A contiguous field of `dims` (x first) over `data`.
    pub fn new<T: ZfpScalar>(data: &'a [T], dims: &[usize]) -> Option<Self> {
        let raw = describe::<T>(data.as_ptr() as *mut c_void, data.len(), dims, None)?;
        Some(Self { raw, _data: PhantomData })
    }

    /// A strided view with non-negative strides (in elements) from `data[0]`.
    pub fn new_strided<T: ZfpScalar>(data: &'a [T], dims: &[usize], strides: &[isize]) -> Option<Self> {
        let raw = describe::<T>(data.as_ptr() as *mut c_void, data.len(), dims, Some(strides))?;
        Some(Self { raw, _data: PhantomData })
    }

    /// A field over device memory (`hipMalloc`), which the caller keeps
    /// alive and unaliased for the field's lifetime.
    ///
    /// # Safety
    /// `ptr` must point to `dims` elements of `T` laid out contiguously.
    pub unsafe fn from_device<T: ZfpScalar>(ptr: *const T, dims: &[usize]) -> Option<Self> {
        let len = dims.iter().try_fold(1usize, |a, &d| a.checked_mul(d))?;
        let raw = describe::<T>(ptr as *mut c_void, len, dims, None)?;
        Some(Self { raw, _data: PhantomData })
    }
}

impl<'a> ZfpFieldMut<'a> {
    pub fn new<T: ZfpScalar>(data: &'a mut [T], dims: &[usize]) -> Option<Self> {
        let raw = describe::<T>(data.as_mut_ptr() as *mut c_void, data.len(), dims, None)?;
        Some(Self { raw, _data: PhantomData })
    }

    /// Strided destinations, aliasing ones included: those decode in the
    /// serial block order, the last block writing an element winning.
    pub fn new_strided<T: ZfpScalar>(data: &'a mut [T], dims: &[usize], strides: &[isize]) -> Option<Self> {
        let raw = describe::<T>(data.as_mut_ptr() as *mut c_void, data.len(), dims, Some(strides))?;
        Some(Self { raw, _data: PhantomData })
    }

    /// # Safety
    /// As [`ZfpField::from_device`], for writing.
    pub unsafe fn from_device<T: ZfpScalar>(ptr: *mut T, dims: &[usize]) -> Option<Self> {
        let len = dims.iter().try_fold(1usize, |a, &d| a.checked_mul(d))?;
        let raw = describe::<T>(ptr as *mut c_void, len, dims, None)?;
        Some(Self { raw, _data: PhantomData })
    }
}

/// A zfp_stream owned by Rust.
struct Stream(*mut ffi::zfp_stream);

impl Stream {
    fn scratch(c: &ZfpConfig) -> Self {
        let z = unsafe { ffi::zfp_stream_open(core::ptr::null_mut()) };
        assert!(!z.is_null(), "zfp_stream_open");
        let s = Stream(z);
        s.configure(c);
        s
    }

    fn configure(&self, c: &ZfpConfig) {
        unsafe {
            ffi::zfp_stream_set_params(self.0, c.min_bits, c.max_bits, c.max_prec, c.min_exp);
            ffi::zfp_stream_set_rounding(self.0, c.rounding.c_mode(), c.rounding.tight_error() as i32);
        }
    }
}

impl Drop for Stream {
    fn drop(&mut self) {
        unsafe { ffi::zfp_stream_close(self.0) }
    }
}

/// A bit stream over caller-owned 64-bit words (host or device memory),
/// with zfp-rs's cursor (`src/bitstream/core.rs`).
pub struct HipBitStream<'a> {
    bs: *mut ffi::bitstream,
    z: Stream,
    capacity: usize,
    _words: PhantomData<&'a mut [u64]>,
}

impl<'a> HipBitStream<'a> {
    pub fn new(words: &'a mut [u64]) -> Option<Self> {
        // SAFETY: the slice outlives the stream
        unsafe { Self::from_raw(words.as_mut_ptr(), words.len()) }
    }

    /// A stream over device words (`hipMalloc`).
    ///
    /// # Safety
    /// `words` must point to `len` writable 64-bit words for `'a`.
    pub unsafe fn from_raw(words: *mut u64, len: usize) -> Option<Self> {
        let bytes = len.checked_mul(8)?;
        let bs = ffi::stream_open(words as *mut c_void, bytes);
        if bs.is_null() {
            return None;
        }
        let z = ffi::zfp_stream_open(bs);
        if z.is_null() {
            ffi::stream_close(bs);
            return None;
        }
        Some(Self { bs, z: Stream(z), capacity: bytes, _words: PhantomData })
    }

    /// Run the GPU work of this stream on a HIP stream (`hipStream_t`).
    pub fn set_hip_stream(&mut self, hip_stream: *mut c_void) -> bool {
        unsafe { ffi::zfp_stream_set_hip_stream(self.z.0, hip_stream) != 0 }
    }

    pub fn seek_read(&mut self, offset: u64) {
        unsafe { ffi::stream_rseek(self.bs, offset) }
    }
    pub fn seek_write(&mut self, offset: u64) {
        unsafe { ffi::stream_wseek(self.bs, offset) }
    }
    pub fn read_pos(&self) -> u64 {
        unsafe { ffi::stream_rtell(self.bs) }
    }
    pub fn write_pos(&self) -> u64 {
        unsafe { ffi::stream_wtell(self.bs) }
    }

    /// Compress `field` at the write cursor; the stream size in bytes after
    /// the flush (`ZfpBitStream::compress`, `src/bitstream/ops.rs:218`).
    pub fn compress(&mut self, config: &ZfpConfig, field: &ZfpField<'_>) -> Result<usize, ZfpCompressionError> {
        self.z.configure(config);
        let before = self.write_pos();
        let n = unsafe { ffi::zfp_compress(self.z.0, &field.raw) };
        if n == 0 {
            // zfp_compress gives 0 for a stream that overflowed (its cursor
            // then sits past the capacity) and for a call that wrote nothing
            let after = self.write_pos();
            if after != before {
                let required = usize::try_from((after + 63) / 64 * 8).unwrap_or(usize::MAX);
                return Err(ZfpCompressionError::BufferTooSmall { required, capacity: self.capacity });
            }
            return Err(ZfpCompressionError::Field);
        }
        Ok(n)
    }

    /// Decompress into `field` from the read cursor; the read position in
    /// bytes afterwards, word aligned (`ZfpBitStream::decompress`,
    /// `src/bitstream/ops.rs:98`).  A stream compressed by this library
    /// keeps an offset index beside it, which parallel variable-rate decodes
    /// use; without one the block lengths are read on the host.
    pub fn decompress(&mut self, config: &ZfpConfig, field: &mut ZfpFieldMut<'_>) -> Result<usize, ZfpDecompressionError> {
        self.z.configure(config);
        let before = self.read_pos();
        let n = unsafe { ffi::zfp_decompress(self.z.0, &mut field.raw) };
        if n == 0 {
            // zfp_decompress reports a truncated stream as 0 with the cursor
            // moved past the blocks; a refused field leaves it where it was
            return Err(if self.read_pos() != before {
                ZfpDecompressionError::Truncated
            } else {
                ZfpDecompressionError::Field
            });
        }
        Ok(n)
    }

    /// The offset index kept beside a variable-rate stream: per-block bit
    /// lengths and the start bit of block 0 (a side-car for storage).
    pub fn index(&self) -> Option<(Vec<u32>, u64)> {
        unsafe {
            let mut start = 0u64;
            let n = ffi::zfp_stream_index_export(self.bs, core::ptr::null_mut(), 0, &mut start);
            if n == 0 {
                return None;
            }
            let mut v = vec![0u32; n];
            let m = ffi::zfp_stream_index_export(self.bs, v.as_mut_ptr(), n, &mut start);
            (m == n).then_some((v, start))
        }
    }

    /// Attach an index saved from [`index`][Self::index].
    pub fn set_index(&mut self, lengths: &[u32], start: u64) -> bool {
        unsafe { ffi::zfp_stream_index_import(self.bs, lengths.as_ptr(), lengths.len(), start) != 0 }
    }
}

impl Drop for HipBitStream<'_> {
    fn drop(&mut self) {
        unsafe {
            (*self.z.0).stream = core::ptr::null_mut();
            ffi::stream_close(self.bs);
        }
    }
}
