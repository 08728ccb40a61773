// Barrel of the drop-in modules (ref: src/index.js:13-21, hot-path part):
// the same export names as the reference for fft, field, propagate, metrics,
// doe (phase retrieval), projector, color and png, backed by the MI355X
// library.
export * from './fft.js';
export * from './field.js';
export * from './propagate.js';
export * from './metrics.js';
export * from './doe.js';
export * from './projector.js';
export * from './color.js';
export * from './png.js';
export { setPrecision, getPrecision } from './native.js';
