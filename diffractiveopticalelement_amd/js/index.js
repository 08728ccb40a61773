// Barrel of the drop-in modules (ref: src/index.js:13-21, hot-path part):
// the reference's export names for fft, field, propagate, doe (phase
// retrieval), projector, color and png, backed by the MI355X library.  The
// host-array metric helpers of src/metrics.js are not part of the drop-in
// (screenMetrics reduces on the device); the reference's own module serves
// them.
export * from './fft.js';
export * from './field.js';
export * from './propagate.js';
export * from './doe.js';
export * from './projector.js';
export * from './color.js';
export * from './png.js';
export { setPrecision, getPrecision } from './native.js';
